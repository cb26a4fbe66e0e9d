"""Copy the reference's power-flow validation cases into tests/golden/power_flow/ and its state-estimation
cases into tests/golden/state_estimation/.

Run once in the build container (the reference tree is not present on the GPU box):
    python tests/golden/collect_fixtures.py /root/reference
A case is taken when its params.json lists a method this build runs (newton_raphson, linear,
iterative_current, linear_current) and every component it uses is one this build supports; cases that
expect an error are taken when they exercise the tap changer or an unknown update id. The data files (input, update_batch, *_output*.json, params.json) are copied unchanged
together with their .license sidecars (MPL-2.0).
"""

import json
import shutil
import sys
from pathlib import Path

SUPPORTED = {"node", "line", "link", "generic_branch", "transformer", "shunt", "source", "sym_gen", "asym_gen", "sym_load",
             "asym_load", "transformer_tap_regulator", "fault"}
FILES = ["input.json", "update_batch.json", "sym_output.json", "asym_output.json", "sym_output_batch.json",
         "asym_output_batch.json", "params.json"]


def components(path):
    d = json.load(open(path))["data"]
    comps = set()
    for sc in d if isinstance(d, list) else [d]:
        comps |= set(sc)
    return comps


def main(ref_root):
    src_root = Path(ref_root) / "tests" / "data" / "power_flow"
    dst_root = Path(__file__).resolve().parent / "power_flow"
    taken = []
    for params in sorted(src_root.rglob("params.json")):
        case = params.parent
        p = json.load(open(params))
        methods = p.get("calculation_method")
        methods = methods if isinstance(methods, list) else [methods]
        if not ({"newton_raphson", "linear", "iterative_current", "linear_current"} & set(methods)):
            continue
        # expected failures are kept for the tap changer (its iteration limits) and for update ids that do not
        # exist
        if "raises" in p and "tap_changing_strategy" not in p and p["raises"].get("raises") != "IDNotFound":
            continue
        # cases the reference itself does not match (xfail with an AssertionError) pin nothing
        if p.get("xfail", {}).get("raises") == "AssertionError":
            continue
        comps = set()
        for f in FILES[:-1]:
            if (case / f).exists():
                comps |= components(case / f)
        if not comps <= SUPPORTED:
            continue
        rel = case.relative_to(src_root)
        dst = dst_root / rel
        dst.mkdir(parents=True, exist_ok=True)
        for f in FILES:
            if (case / f).exists():
                shutil.copy(case / f, dst / f)
                if (case / (f + ".license")).exists():
                    shutil.copy(case / (f + ".license"), dst / (f + ".license"))
        taken.append(str(rel))
    print("\n".join(taken))


SE_SUPPORTED = {"node", "line", "link", "generic_branch", "transformer", "shunt", "source", "sym_gen", "asym_gen",
                "sym_load", "asym_load", "sym_voltage_sensor", "sym_power_sensor"}


def main_state_estimation(ref_root):
    """State-estimation cases with a symmetric expected output whose params.json lists newton_raphson and whose
    components this build takes (symmetric sensors, no current sensors or three-winding transformers), into
    tests/golden/state_estimation/."""
    src_root = Path(ref_root) / "tests" / "data" / "state_estimation"
    dst_root = Path(__file__).resolve().parent / "state_estimation"
    taken = []
    for params in sorted(src_root.rglob("params.json")):
        case = params.parent
        p = json.load(open(params))
        methods = p.get("calculation_method")
        methods = methods if isinstance(methods, list) else [methods]
        if "newton_raphson" not in methods or "raises" in p or "xfail" in p:
            continue
        if not ((case / "sym_output.json").exists() or (case / "sym_output_batch.json").exists()):
            continue
        comps = set()
        for f in ("input.json", "update_batch.json", "sym_output.json", "sym_output_batch.json"):
            if (case / f).exists():
                comps |= components(case / f)
        if not comps <= SE_SUPPORTED:
            continue
        rel = case.relative_to(src_root)
        dst = dst_root / rel
        dst.mkdir(parents=True, exist_ok=True)
        for f in ("input.json", "update_batch.json", "sym_output.json", "sym_output_batch.json", "params.json"):
            if (case / f).exists():
                shutil.copy(case / f, dst / f)
                if (case / (f + ".license")).exists():
                    shutil.copy(case / (f + ".license"), dst / (f + ".license"))
        taken.append(str(rel))
    print("\n".join(taken))


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    main(root)
    main_state_estimation(root)
