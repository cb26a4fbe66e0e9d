"""Helpers to run the reference's power-flow validation cases (tests/golden/power_flow/**).

The comparison follows the reference's own rule (tests/unit/utils.py compare_result): every attribute
present in the expected file is compared with numpy.allclose(rtol, atol) where atol may be a dict of
attribute-name patterns; u_angle is compared after removing the global angle of the first entry and
angles are compared as complex phasors together with their magnitude.
"""

from __future__ import annotations

import json
import re
from pathlib import Path

import numpy as np

from pgm_formats import batch_to_uniform, load_dataset

GOLDEN = Path(__file__).resolve().parent / "golden" / "power_flow"
GOLDEN_SE = Path(__file__).resolve().parent / "golden" / "state_estimation"


def list_cases(method: str = "newton_raphson", tap: bool | None = None, raises: bool = False, root: Path = GOLDEN):
    """(case, symmetric, batch) of every fixture whose params.json lists the calculation method.

    tap: None for every case, False without / True only with a tap changing strategy. raises: the cases
    that expect an error instead of the ones with outputs (params "raises", or "xfail" naming the error
    the reference's own run raises)."""
    cases = []
    for params in sorted(root.rglob("params.json")):
        case = params.parent
        p = json.load(open(params))
        methods = p.get("calculation_method")
        methods = methods if isinstance(methods, list) else [methods]
        if method not in methods:
            continue
        if tap is not None and tap != ("tap_changing_strategy" in p):
            continue
        if raises != ("raises" in p or "xfail" in p):
            continue
        if raises:
            cases.append((str(case.relative_to(root)), True, (case / "update_batch.json").exists()))
            continue
        for sym, stem in ((True, "sym_output"), (False, "asym_output")):
            for batch in (False, True):
                f = case / (stem + ("_batch" if batch else "") + ".json")
                if f.exists():
                    cases.append((str(case.relative_to(root)), sym, batch))
    return cases


def load_case(name: str, sym: bool, batch: bool, root: Path = GOLDEN):
    case = root / name
    params = json.load(open(case / "params.json"))
    input_data = load_dataset(case / "input.json")
    stem = ("sym_output" if sym else "asym_output") + ("_batch" if batch else "")
    expected = load_dataset(case / (stem + ".json")) if (case / (stem + ".json")).exists() else None
    update = None
    n_scenarios = 1
    if batch:
        raw = load_dataset(case / "update_batch.json")
        n_scenarios = len(raw)
        update = batch_to_uniform(raw, input_data)
        expected = expected if isinstance(expected, list) else [expected]
    else:
        expected = [expected]
    return params, input_data, update, n_scenarios, expected


def compare_result(actual: dict, expected: dict, rtol: float, atol, scenario: int = 0):
    """actual[comp] has shape (n_scenarios, n); expected[comp] is one scenario's records."""
    for key, exp in expected.items():
        act = actual[key][scenario]
        # expected rows may be a subset / permutation: match by id
        if np.all(exp["id"] == np.iinfo("i4").min):
            act = act[: len(exp)]  # rows without id: positional
        else:
            idx = {int(i): k for k, i in enumerate(act["id"])}
            sel = np.array([idx[int(i)] for i in exp["id"]], dtype=np.int64)
            act = act[sel]
        for col in exp.dtype.names:
            if col == "id":
                continue
            e = exp[col]
            a = act[col]
            if e.dtype == np.float64:
                if np.all(np.isnan(e)):
                    continue
            elif e.dtype == np.int8:
                if np.all(e == np.iinfo("i1").min):
                    continue
            elif e.dtype == np.int32:
                if np.all(e == np.iinfo("i4").min):
                    continue
            if e.ndim == 1 and a.ndim == 2:
                if col == "u_angle":
                    e = np.stack((e, e - 2.0 / 3.0 * np.pi, e + 2.0 / 3.0 * np.pi), axis=-1)
                    e = np.angle(np.exp(1j * e))
                else:
                    e = e.reshape(-1, 1)
            if col == "u_angle":
                a = a - a.ravel()[0]
                e = e - e.ravel()[0]
            if col.endswith("_angle"):
                mag = col[: -len("_angle")]
                if np.all(np.isnan(exp[mag])):
                    continue
                a = act[mag] * np.exp(1j * a)
                em = exp[mag]
                if em.ndim == 1 and a.ndim == 2:
                    em = em.reshape(-1, 1)
                e = em * np.exp(1j * e)
            tol = atol
            if isinstance(atol, dict):
                tol = atol["default"]
                for pattern, t in atol.items():
                    if re.fullmatch(pattern, col):
                        tol = t
                        break
            assert np.allclose(a, e, rtol=rtol, atol=tol), (
                f"{key}.{col} scenario {scenario}: actual {a} expected {e} (rtol={rtol}, atol={tol})"
            )
