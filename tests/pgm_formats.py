"""Dataset record layouts and JSON fixture loading for the tests.

The layouts are the row-based dataset layouts of power-grid-model
(code_generation/data/attribute_classes/{input,update,output}.json, structs in
power_grid_model_c/power_grid_model/include/power_grid_model/auxiliary/{input,update,output}.hpp),
written out here with C alignment so that tests can check the product's exported metadata
(PGM_meta_*) against an independent statement of the same layout.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

NAN_INT8 = np.int8(-128)
NAN_INT32 = np.int32(np.iinfo(np.int32).min)

_D = "f8"
_D3 = ("f8", (3,))


def _struct(fields):
    """Build a numpy dtype with C struct alignment from (name, format) pairs."""
    names, formats, offsets = [], [], []
    off = 0
    max_align = 1
    for name, fmt in fields:
        dt = np.dtype(fmt)
        align = dt.base.alignment
        max_align = max(max_align, align)
        off = (off + align - 1) // align * align
        names.append(name)
        formats.append(fmt)
        offsets.append(off)
        off += dt.itemsize
    size = (off + max_align - 1) // max_align * max_align
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size}, align=True)


_BASE = [("id", "i4")]
_BRANCH_IN = _BASE + [("from_node", "i4"), ("to_node", "i4"), ("from_status", "i1"), ("to_status", "i1")]
_APPL_IN = _BASE + [("node", "i4"), ("status", "i1")]

INPUT = {
    "node": _struct(_BASE + [("u_rated", _D)]),
    "line": _struct(
        _BRANCH_IN
        + [(n, _D) for n in ("r1", "x1", "c1", "tan1", "r0", "x0", "c0", "tan0", "i_n")]
    ),
    "link": _struct(_BRANCH_IN),
    "generic_branch": _struct(_BRANCH_IN + [(n, _D) for n in ("r1", "x1", "g1", "b1", "k", "theta", "sn")]),
    "transformer": _struct(
        _BRANCH_IN
        + [(n, _D) for n in ("u1", "u2", "sn", "uk", "pk", "i0", "p0", "i0_zero_sequence", "p0_zero_sequence")]
        + [(n, "i1") for n in ("winding_from", "winding_to", "clock", "tap_side", "tap_pos", "tap_min", "tap_max",
                               "tap_nom")]
        + [(n, _D) for n in ("tap_size", "uk_min", "uk_max", "pk_min", "pk_max", "r_grounding_from",
                              "x_grounding_from", "r_grounding_to", "x_grounding_to")]
    ),
    "shunt": _struct(_APPL_IN + [(n, _D) for n in ("g1", "b1", "g0", "b0")]),
    "source": _struct(_APPL_IN + [(n, _D) for n in ("u_ref", "u_ref_angle", "sk", "rx_ratio", "z01_ratio")]),
    "sym_gen": _struct(_APPL_IN + [("type", "i1"), ("p_specified", _D), ("q_specified", _D)]),
    "asym_gen": _struct(_APPL_IN + [("type", "i1"), ("p_specified", _D3), ("q_specified", _D3)]),
    "sym_load": _struct(_APPL_IN + [("type", "i1"), ("p_specified", _D), ("q_specified", _D)]),
    "asym_load": _struct(_APPL_IN + [("type", "i1"), ("p_specified", _D3), ("q_specified", _D3)]),
    "fault": _struct(_BASE + [("status", "i1"), ("fault_type", "i1"), ("fault_phase", "i1"), ("fault_object", "i4"),
                              ("r_f", _D), ("x_f", _D)]),
    "transformer_tap_regulator": _struct(
        _BASE + [("regulated_object", "i4"), ("status", "i1"), ("control_side", "i1")]
        + [(n, _D) for n in ("u_set", "u_band", "line_drop_compensation_r", "line_drop_compensation_x")]
    ),
    "sym_voltage_sensor": _struct(_BASE + [("measured_object", "i4")]
                                  + [(n, _D) for n in ("u_sigma", "u_measured", "u_angle_measured")]),
    "sym_power_sensor": _struct(_BASE + [("measured_object", "i4"), ("measured_terminal_type", "i1")]
                                + [(n, _D) for n in ("power_sigma", "p_measured", "q_measured", "p_sigma",
                                                     "q_sigma")]),
}

UPDATE = {
    "line": _struct(_BASE + [("from_status", "i1"), ("to_status", "i1")]),
    "link": _struct(_BASE + [("from_status", "i1"), ("to_status", "i1")]),
    "generic_branch": _struct(_BASE + [("from_status", "i1"), ("to_status", "i1")]),
    "transformer": _struct(_BASE + [("from_status", "i1"), ("to_status", "i1"), ("tap_pos", "i1")]),
    "shunt": _struct(_BASE + [("status", "i1")] + [(n, _D) for n in ("g1", "b1", "g0", "b0")]),
    "source": _struct(_BASE + [("status", "i1")] + [(n, _D) for n in ("u_ref", "u_ref_angle", "sk", "rx_ratio",
                                                                      "z01_ratio")]),
    "sym_gen": _struct(_BASE + [("status", "i1"), ("p_specified", _D), ("q_specified", _D)]),
    "asym_gen": _struct(_BASE + [("status", "i1"), ("p_specified", _D3), ("q_specified", _D3)]),
    "sym_load": _struct(_BASE + [("status", "i1"), ("p_specified", _D), ("q_specified", _D)]),
    "asym_load": _struct(_BASE + [("status", "i1"), ("p_specified", _D3), ("q_specified", _D3)]),
    "transformer_tap_regulator": _struct(
        _BASE + [("status", "i1")]
        + [(n, _D) for n in ("u_set", "u_band", "line_drop_compensation_r", "line_drop_compensation_x")]
    ),
    "sym_voltage_sensor": _struct(_BASE + [(n, _D) for n in ("u_sigma", "u_measured", "u_angle_measured")]),
    "sym_power_sensor": _struct(_BASE + [(n, _D) for n in ("power_sigma", "p_measured", "q_measured", "p_sigma",
                                                           "q_sigma")]),
}


def _output(v):
    base = _BASE + [("energized", "i1")]
    node = _struct(base + [(n, v) for n in ("u_pu", "u", "u_angle", "p", "q")])
    branch = _struct(base + [("loading", _D)] + [(n, v) for n in ("p_from", "q_from", "i_from", "s_from", "p_to",
                                                                  "q_to", "i_to", "s_to")])
    appl = _struct(base + [(n, v) for n in ("p", "q", "i", "s", "pf")])
    return {
        "node": node,
        "line": branch,
        "link": branch,
        "generic_branch": branch,
        "transformer": branch,
        "shunt": appl,
        "source": appl,
        "sym_gen": appl,
        "asym_gen": appl,
        "sym_load": appl,
        "asym_load": appl,
        "transformer_tap_regulator": _struct(base + [("tap_pos", "i1")]),
        "fault": _struct(base),
        "sym_voltage_sensor": _struct(base + [("u_residual", v), ("u_angle_residual", v)]),
        "sym_power_sensor": _struct(base + [("p_residual", v), ("q_residual", v)]),
    }


SYM_OUTPUT = _output(_D)
ASYM_OUTPUT = _output(_D3)

DATASETS = {"input": INPUT, "update": UPDATE, "sym_output": SYM_OUTPUT, "asym_output": ASYM_OUTPUT}

# component order of the reference's container (all_components.hpp:37-40), the link and the generic branch last
# (the oracle's slots)
COMPONENT_ORDER = ["node", "line", "transformer", "shunt", "source", "sym_gen", "asym_gen", "sym_load", "asym_load",
                   "link", "generic_branch"]
# with the components outside the math model (the tap changer's regulators), in the same container order
ALL_COMPONENTS = COMPONENT_ORDER + ["transformer_tap_regulator", "fault", "sym_voltage_sensor", "sym_power_sensor"]


def null_value(fmt: np.dtype):
    base = fmt.base
    if base.kind == "f":
        return np.nan
    if base == np.int8:
        return NAN_INT8
    if base == np.int32:
        return NAN_INT32
    raise TypeError(fmt)


def empty(dtype: np.dtype, n: int) -> np.ndarray:
    """Array of n records with every attribute null (NaN / na), like initialize_array."""
    arr = np.zeros(n, dtype=dtype)
    for name in dtype.names:
        arr[name] = null_value(dtype.fields[name][0])
    return arr


def rows_to_array(rows: list, dtype: np.dtype, attributes: list[str] | None = None) -> np.ndarray:
    """Rows are dicts, or lists of values in the order given by the file's "attributes" entry."""
    arr = empty(dtype, len(rows))
    for i, row in enumerate(rows):
        items = row.items() if isinstance(row, dict) else zip(attributes, row)
        for k, v in items:
            if k in dtype.names:
                arr[k][i] = v
    return arr


def load_dataset(path: Path, dataset_type: str | None = None):
    """Load a reference JSON dataset file into {component: array} (single) or a list of those (batch)."""
    with open(path) as f:
        d = json.load(f)
    dtypes = DATASETS[dataset_type or d["type"]]
    attrs = d.get("attributes", {})
    if d["is_batch"]:
        return [{k: rows_to_array(v, dtypes[k], attrs.get(k)) for k, v in sc.items() if k in dtypes}
                for sc in d["data"]]
    return {k: rows_to_array(v, dtypes[k], attrs.get(k)) for k, v in d["data"].items() if k in dtypes}


def unsupported_components(path: Path) -> set[str]:
    with open(path) as f:
        d = json.load(f)
    data = d["data"]
    comps = set()
    for sc in data if isinstance(data, list) else [data]:
        comps |= set(sc.keys())
    return comps - set(ALL_COMPONENTS)


def batch_to_uniform(batch: list[dict], input_data: dict) -> dict:
    """Turn a (possibly sparse) update batch into a uniform one by padding each scenario with no-op rows.

    A padding row carries the id of a component that scenario does not touch and null attributes, which
    the reference's update semantics treat as "keep the current value".
    """
    comps = sorted({c for sc in batch for c in sc}, key=ALL_COMPONENTS.index)
    out = {}
    for comp in comps:
        dtype = UPDATE[comp]
        n_max = max(len(sc.get(comp, [])) for sc in batch)
        all_ids = input_data[comp]["id"]
        arr = empty(dtype, n_max * len(batch)).reshape(len(batch), n_max)
        for s, sc in enumerate(batch):
            rows = sc.get(comp, empty(dtype, 0))
            arr[s, : len(rows)] = rows
            if len(rows) < n_max:
                used = set(int(x) for x in rows["id"])
                free = [int(x) for x in all_ids if int(x) not in used]
                if len(free) < n_max - len(rows):
                    raise ValueError("cannot pad update batch")
                for j in range(len(rows), n_max):
                    arr[s, j]["id"] = free[j - len(rows)]
        out[comp] = arr
    return out
