"""Host-side checks of the product library, no GPU needed.

* the C ABI exports every entry point include/power_grid_model_amd.h declares;
* the dataset metadata the Python layer builds its dtypes from has the reference's layouts
  (power_grid_model_c/include/power_grid_model_c/meta_data.h, attribute_classes/*.json), here the
  C-aligned layouts restated in pgm_formats.py;
* the product's bus numbering, minimum-degree ordering and fill-ins match the reference's known answers
  (test_sparse_ordering.cpp:21-35, test_topology.cpp:380-469) and the oracle bit for bit on random grids;
* the device factorisation schedule respects every dependency of the block LU.
"""

from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as ol
import pgm_formats as pf

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "power-grid-model_amd" / "power_grid_model_amd" / "_lib" / "libpower_grid_model_c.so"
if not LIB.exists():
    subprocess.run([sys.executable, str(ROOT / "power-grid-model_amd" / "build.py")], check=True)

from power_grid_model_amd import (  # noqa: E402
    PowerGridModel,
    initialize_array,
    inspect_topology,
    minimum_degree_ordering,
    plan_stats,
    power_grid_meta_data,
    synthetic,
)
from power_grid_model_amd._lib import lib  # noqa: E402
import power_grid_model_amd as pgm_errors  # noqa: E402


def declared_symbols():
    text = (ROOT / "include" / "power_grid_model_amd.h").read_text()
    return sorted(set(re.findall(r"PGM_AMD_API[^;(]*?\b(PGM_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


# every C function the reference's Python layer binds at import
# (src/power_grid_model/_core/power_grid_core.py, @make_c_binding methods); restated as data
REFERENCE_PYTHON_BINDINGS = [
    "version", "create_dataset_const_from_mutable", "create_handle", "error_code", "error_message", "n_failed_scenarios", "failed_scenarios",
    "batch_errors", "clear_error", "meta_n_datasets", "meta_get_dataset_by_idx", "meta_dataset_name",
    "meta_n_components", "meta_get_component_by_idx", "meta_component_name", "meta_component_size",
    "meta_component_alignment", "meta_n_attributes", "meta_get_attribute_by_idx", "meta_attribute_name",
    "meta_attribute_ctype", "meta_attribute_offset", "is_little_endian", "create_options", "destroy_options",
    "set_calculation_type", "set_calculation_method", "set_symmetric", "set_err_tol", "set_max_iter",
    "set_threading", "set_tap_changing_strategy", "set_short_circuit_voltage_scaling",
    "set_experimental_features", "create_model", "update_model", "copy_model", "get_indexer", "destroy_model",
    "calculate", "dataset_info_name", "dataset_info_is_batch", "dataset_info_batch_size",
    "dataset_info_n_components", "dataset_info_component_name", "dataset_info_elements_per_scenario",
    "dataset_info_total_elements", "dataset_info_has_attribute_indications", "dataset_info_n_attribute_indications",
    "dataset_info_attribute_name", "create_dataset_const", "destroy_dataset_const", "dataset_const_add_buffer",
    "dataset_const_add_attribute_buffer", "dataset_const_set_next_cartesian_product_dimension",
    "dataset_const_get_info", "dataset_writable_get_info", "dataset_writable_set_buffer",
    "dataset_writable_set_attribute_buffer", "create_dataset_mutable", "destroy_dataset_mutable",
    "dataset_mutable_add_buffer", "dataset_mutable_add_attribute_buffer", "dataset_mutable_get_info",
    "create_deserializer_from_binary_buffer", "create_deserializer_from_null_terminated_string",
    "deserializer_get_dataset", "deserializer_parse_to_buffer", "destroy_deserializer", "create_serializer",
    "serializer_get_to_binary_buffer", "serializer_get_to_zero_terminated_string", "destroy_serializer",
]


def test_exports_everything_the_reference_python_layer_binds():
    missing = [n for n in REFERENCE_PYTHON_BINDINGS if not hasattr(lib, "PGM_" + n)]
    assert missing == []


def test_serialization_fails_with_a_regular_error():
    import ctypes as C

    from power_grid_model_amd._lib import Handle

    h = Handle()
    f = lib.PGM_create_serializer
    f.restype = C.c_void_p
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    assert f(h.ptr, None, 0) is None
    assert h.code() == 1 and "not supported" in h.message()


def test_meta_layouts_match_reference_formats():
    for dataset, comps in pf.DATASETS.items():
        meta = power_grid_meta_data[dataset]
        for comp, dtype in comps.items():
            assert comp in meta, (dataset, comp)
            got = meta[comp].dtype
            assert got.itemsize == dtype.itemsize, (dataset, comp)
            assert got.alignment == dtype.alignment, (dataset, comp)
            for name in dtype.names:
                assert got.fields[name][1] == dtype.fields[name][1], (dataset, comp, name)
                assert got.fields[name][0] == dtype.fields[name][0], (dataset, comp, name)


def test_initialize_array_fills_null_values():
    a = initialize_array("input", "sym_load", 3)
    assert (a["id"] == np.iinfo(np.int32).min).all()
    assert (a["status"] == np.iinfo(np.int8).min).all()
    assert np.isnan(a["p_specified"]).all()


# ------------------------------------------------------------------------------------------------
# ordering and topology: reference known answers
# ------------------------------------------------------------------------------------------------
def test_minimum_degree_reference_vector():
    graph = {0: [3, 5], 1: [4, 5, 8], 2: [4, 5, 6], 3: [6, 7], 4: [6, 8], 6: [7, 8, 9], 7: [8, 9], 8: [9]}
    alpha, fills = minimum_degree_ordering(graph)
    assert alpha == list(range(10))
    assert fills == [(3, 5), (4, 5), (5, 8), (5, 6), (5, 7)]


def _grid(n_node, branches, sources, statuses=None):
    node = initialize_array("input", "node", n_node)
    node["id"] = np.arange(n_node)
    node["u_rated"] = 10.5e3
    line = initialize_array("input", "line", len(branches))
    line["id"] = 1000 + np.arange(len(branches))
    line["from_node"] = [b[0] for b in branches]
    line["to_node"] = [b[1] for b in branches]
    st = np.ones((len(branches), 2), dtype=np.int8) if statuses is None else np.array(statuses, dtype=np.int8)
    line["from_status"] = st[:, 0]
    line["to_status"] = st[:, 1]
    line["r1"] = 0.1
    line["x1"] = 0.05
    line["c1"] = 1e-8
    line["tan1"] = 0.0
    line["r0"] = 0.3
    line["x0"] = 0.15
    line["c0"] = 1e-8
    line["tan0"] = 0.0
    line["i_n"] = 500.0
    src = initialize_array("input", "source", len(sources))
    src["id"] = 5000 + np.arange(len(sources))
    src["node"] = sources
    src["status"] = 1
    src["u_ref"] = 1.0
    src["sk"] = 1e10
    return {"node": node, "line": line, "source": src}


def test_cycle_reorder_9_nodes():
    br = [(0, 3), (0, 5), (1, 4), (1, 5), (1, 8), (2, 4), (2, 5), (2, 6), (3, 6), (3, 7), (4, 6), (4, 8), (6, 7),
          (6, 8), (6, 9), (7, 8), (7, 9), (8, 9)]
    group, pos, fills = inspect_topology(PowerGridModel(_grid(10, br, [0])), 10)
    assert group.tolist() == [0] * 10
    assert pos.tolist() == list(range(10))
    assert [f[1:] for f in fills] == [(3, 5), (4, 5), (5, 8), (5, 6), (5, 7)]


def test_cycle_reorder_7_nodes():
    br = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5), (1, 4), (6, 0), (6, 2), (5, 1), (3, 1), (6, 1), (2, 1)]
    group, pos, fills = inspect_topology(PowerGridModel(_grid(7, br, [0])), 7)
    assert pos.tolist() == [0, 3, 1, 2, 4, 5, 6]
    assert [f[1:] for f in fills] == [(5, 6), (2, 6), (4, 6)]


def test_radial_reversed_dfs_and_islands():
    group, pos, fills = inspect_topology(PowerGridModel(_grid(4, [(0, 1), (1, 2), (1, 3)], [0])), 4)
    assert pos.tolist() == [3, 2, 1, 0] and fills == []
    group, pos, _ = inspect_topology(PowerGridModel(_grid(4, [(0, 1), (1, 2)], [0, 3], [(1, 1), (1, 0)])), 4)
    assert group.tolist() == [0, 0, -1, 1]
    assert pos[2] == -1 and pos[3] == 0


@pytest.mark.parametrize("seed", range(6))
def test_topology_matches_oracle_on_random_grids(seed):
    """Random meshed grids with open sides and several sources: node groups, bus positions and fill-ins
    equal the oracle's bit for bit."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(30, 120))
    parent = [int(rng.integers(max(0, i - 5), i)) for i in range(1, n)]
    br = [(p, i) for i, p in zip(range(1, n), parent)]
    for _ in range(int(rng.integers(2, 25))):
        a, b = rng.choice(n, 2, replace=False)
        br.append((int(a), int(b)))
    st = [(1, 1)] * len(br)
    for k in rng.choice(len(br), 3, replace=False):
        st[k] = (int(rng.integers(0, 2)), int(rng.integers(0, 2)))
    sources = [0] + ([int(rng.integers(1, n))] if seed % 2 else [])
    group, pos, fills = inspect_topology(PowerGridModel(_grid(n, br, sources, st)), n)
    og, op, nmm, ofills = ol.build_topology(n, br, sources, statuses=st)
    assert group.tolist() == og
    assert pos.tolist() == op
    assert sorted(fills) == sorted(ofills)


# ------------------------------------------------------------------------------------------------
# device schedule
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("make", [lambda: synthetic.radial_200(n_scenarios=1),
                                  lambda: synthetic.meshed_1000(n_scenarios=1),
                                  lambda: synthetic.feeders_64(n_scenarios=1, n_feeder=8)],
                         ids=["radial_200", "meshed_1000", "feeders_8"])
@pytest.mark.parametrize("target", [4, 64, 256])
@pytest.mark.parametrize("packed_lanes", [0, 1, 3, 32])
def test_schedule_respects_dependencies(make, target, packed_lanes):
    """The schedule respects every dependency of the block LU; the interleaved layout of packed groups
    (1, 3, 32 scenarios) keeps rows, diagonals, transposes and Schur triples consistent."""
    case = make()
    st = plan_stats(PowerGridModel(case.input), True, target, packed_lanes)
    assert st["valid"] == 1
    assert st["tasks"] >= 1 and st["levels"] >= 1
    assert st["critical_path"] <= st["buses"]


def test_interleaved_layout_of_switching_topologies():
    """Packed layout of year-long switching topologies (ties closed, feeder lines open, cyclic parts)."""
    case = synthetic.year_20k(n_scenarios=200)
    ind, data = case.update["line"]["indptr"], case.update["line"]["data"]
    checked = 0
    for s in np.flatnonzero(np.diff(ind) > 0)[:4]:
        inp = {k: v.copy() for k, v in case.input.items()}
        line = inp["line"]
        for r in data[ind[s]:ind[s + 1]]:
            k = np.flatnonzero(line["id"] == r["id"])[0]
            line["from_status"][k] = r["from_status"]
            line["to_status"][k] = r["to_status"]
        st = plan_stats(PowerGridModel(inp), True, 256, 1)
        assert st["valid"] == 1
        checked += 1
    assert checked == 4


# ------------------------------------------------------------------------------------------------
# get_indexer over the reference's validation inputs (tests/unit/test_0Z_model_validation.py:150-157:
# shuffled ids of every input component map back to their rows)
# ------------------------------------------------------------------------------------------------
def _golden_inputs():
    from validation_utils import GOLDEN

    return sorted(str(p.parent.relative_to(GOLDEN)) for p in GOLDEN.rglob("input.json"))


@pytest.mark.parametrize("name", _golden_inputs())
def test_get_indexer_on_reference_inputs(name):
    from validation_utils import GOLDEN

    input_data = pf.load_dataset(GOLDEN / name / "input.json")
    model = PowerGridModel(input_data)
    for comp, arr in input_data.items():
        ids = arr["id"].copy()
        np.random.default_rng(3).shuffle(ids)
        idx = model.get_indexer(comp, ids)
        assert (arr["id"][idx] == ids).all(), comp
    with pytest.raises(pgm_errors.IDNotFound):
        model.get_indexer("node", np.array([-12345], dtype=np.int32))
