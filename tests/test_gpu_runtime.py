"""The runtime around the kernels, on the GPU: graph-captured Newton-Raphson iterations against one launch per
level, update data resident in HBM against uploads per call, and the staged device-to-host copy of large
outputs. Each pair must give the same bytes: the kernels and their arithmetic are the same, only the way the
work reaches the device differs.
"""

import os

import numpy as np
import pytest

pgm = pytest.importorskip("power_grid_model_amd")
from power_grid_model_amd import DeviceBatch, PowerGridModel, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def _subset(case, scenarios):
    upd = {}
    for comp, v in case.update.items():
        if isinstance(v, dict):
            ind = v["indptr"]
            rows = [v["data"][ind[s]:ind[s + 1]] for s in scenarios]
            indptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
            upd[comp] = {"indptr": indptr, "data": np.concatenate(rows) if rows else v["data"][:0]}
        else:
            upd[comp] = v[scenarios]
    return upd


class _Env:
    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = v

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _same(a, b):
    for comp in a:
        assert a[comp].tobytes() == b[comp].tobytes(), comp


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_iterations_equal_level_launches(precision):
    """year_20k sample (plain, union and low-rank lanes on two streams): the captured iteration (default)
    and one launch per level (PGM_AMD_GRAPH=0) give identical outputs and iteration counts, on the first
    run (capture) and on a second run of the same batch (replay)."""
    case = synthetic.year_20k()
    sw = np.flatnonzero(np.diff(case.update["line"]["indptr"]) > 0)
    upd = _subset(case, list(range(0, 40)) + list(sw[:24]))
    model = PowerGridModel(case.input)
    with _Env(PGM_AMD_GRAPH="0"):
        ref_b = DeviceBatch(model, upd, factor_precision=precision)
        ref_b.run()
        ref, ref_it = ref_b.fetch(), ref_b.iterations()
        assert ref_b.stats()["graph_launches"] == 0
    b = DeviceBatch(model, upd, factor_precision=precision)
    for _ in range(2):
        b.run()
        # the captured path ran: one launch per iteration, at most one beyond the slowest scenario's count
        assert ref_it.max() <= b.stats()["graph_launches"] <= ref_it.max() + 1
        _same(b.fetch(), ref)
        assert (b.iterations() == ref_it).all()


def test_graph_iterations_max_iter_and_divergence():
    """Every scenario runs to max_iterations and fails (error tolerance 0): the captured loop launches exactly
    max_iterations iterations and reports the reference's divergence for all of them, as the level launches
    do."""
    case = synthetic.radial_200(n_scenarios=70)
    model = PowerGridModel(case.input)
    outs = []
    for graph in ("0", "1"):
        with _Env(PGM_AMD_GRAPH=graph):
            with pytest.raises(pgm.PowerGridBatchError) as info:
                model.calculate_power_flow(update_data=case.update, error_tolerance=0.0, max_iterations=3)
            outs.append(info.value)
    assert sorted(outs[0].failed_scenarios.tolist()) == list(range(70))
    assert sorted(outs[1].failed_scenarios.tolist()) == list(range(70))
    assert [str(e) for e in outs[0].errors] == [str(e) for e in outs[1].errors]


def test_resident_update_data():
    """Load updates made resident in HBM give the outputs of per-call uploads, also after the host arrays
    are passed again in a second call, and after the resident copies are released."""
    case = synthetic.meshed_1000(n_scenarios=200)
    sel = {"node": ["u_pu", "u_angle"], "line": ["p_from", "q_to"], "sym_load": ["p", "q"]}
    plain = PowerGridModel(case.input).calculate_power_flow(update_data=case.update, output_component_types=sel)
    model = PowerGridModel(case.input)
    n = model.make_resident(case.update)
    assert n == case.update["sym_load"].nbytes
    assert model.make_resident(case.update) == 0  # already resident
    for _ in range(2):
        out = model.calculate_power_flow(update_data=case.update, output_component_types=sel)
        for comp in sel:
            for attr in sel[comp]:
                assert out[comp][attr].tobytes() == plain[comp][attr].tobytes(), (comp, attr)
    b = DeviceBatch(model, case.update)
    b.set_output(sel)
    b.run()
    out = b.fetch()
    for comp in sel:
        for attr in sel[comp]:
            assert out[comp][attr].tobytes() == plain[comp][attr].tobytes(), (comp, attr)
    model.release_device_memory()
    out = model.calculate_power_flow(update_data=case.update, output_component_types=sel)
    assert out["node"]["u_pu"].tobytes() == plain["node"]["u_pu"].tobytes()


def _resident_same(inp, upd, sel):
    plain = PowerGridModel(inp).calculate_power_flow(update_data=upd, output_component_types=sel)
    model = PowerGridModel(inp)
    assert model.make_resident(upd) > 0
    out = model.calculate_power_flow(update_data=upd, output_component_types=sel)
    for comp in sel:
        for attr in sel[comp]:
            assert out[comp][attr].tobytes() == plain[comp][attr].tobytes(), (comp, attr)


def test_resident_update_data_columnar_shuffled_and_positional():
    """The ids of resident update buffers are checked on the device: a columnar buffer, a buffer whose
    scenarios list their loads in different orders (each scenario then resolves its own ids), and one
    without ids (rows by position) all give the outputs of per-call uploads."""
    case = synthetic.meshed_1000(n_scenarios=100)
    sel = {"node": ["u_pu", "u_angle"], "sym_load": ["p", "q"]}
    rows = case.update["sym_load"]
    cols = {f: np.ascontiguousarray(rows[f]) for f in ("id", "status", "p_specified", "q_specified")}
    _resident_same(case.input, {"sym_load": cols}, sel)
    rng = np.random.default_rng(5)
    shuffled = np.stack([r[rng.permutation(len(r))] for r in rows])
    _resident_same(case.input, {"sym_load": shuffled}, sel)
    # positional: every load in input order, ids NA
    pos = np.empty_like(rows)
    order = np.argsort(np.argsort(case.input["sym_load"]["id"]))
    pos[:] = rows[:, np.argsort(rows[0]["id"])][:, order]
    pos["id"] = np.iinfo(np.int32).min
    _resident_same(case.input, {"sym_load": pos}, sel)


def test_staged_copy_of_large_outputs():
    """Whole records of 4096 meshed_1000 scenarios (node 196 MB, line 344 MB) reach the host through the
    staged copy; the first 64 scenarios equal a batch of just those 64 (lanes are independent, so the
    results are the same bytes), whose outputs are small enough for one direct copy."""
    case = synthetic.meshed_1000(n_scenarios=4096)
    model = PowerGridModel(case.input)
    big = DeviceBatch(model, case.update)
    big.run()
    out = big.fetch()
    small = DeviceBatch(model, {"sym_load": case.update["sym_load"][:64]})
    small.run()
    ref = small.fetch()
    for comp in ref:
        assert out[comp][:64].tobytes() == ref[comp].tobytes(), comp
    # and the last 64 through a batch of their own
    tail = DeviceBatch(model, {"sym_load": case.update["sym_load"][-64:]})
    tail.run()
    ref = tail.fetch()
    for comp in ref:
        assert out[comp][-64:].tobytes() == ref[comp].tobytes(), comp


def test_autotuned_schedule_gives_the_same_results():
    """The autotuner builds and runs the batch with candidate schedules and keeps the fastest for the model;
    every schedule computes the same operations in the same order per lane (pull-form Schur updates in
    elimination order), so the outputs after tuning equal the default schedule's byte for byte."""
    case = synthetic.meshed_1000(n_scenarios=300)
    sel = {"node": ["u_pu", "u_angle"], "line": ["p_from", "q_from"]}
    ref = PowerGridModel(case.input).calculate_power_flow(update_data=case.update, output_component_types=sel,
                                                          factor_precision="fp32")
    model = PowerGridModel(case.input)
    res = model.autotune(case.update, factor_precision="fp32")
    assert res["params"]["plain_waves"] in (3072, 6144, 12288, 24576, 49152)
    assert len(res["trials"]) >= 5 and all(ms > 0 for _, ms in res["trials"])
    best = min(ms for _, ms in res["trials"])
    assert any(p == res["params"] and ms == best for p, ms in res["trials"])
    out = model.calculate_power_flow(update_data=case.update, output_component_types=sel, factor_precision="fp32")
    for comp in sel:
        for attr in sel[comp]:
            assert out[comp][attr].tobytes() == ref[comp][attr].tobytes(), (comp, attr)


def test_plans_kept_between_calculations_and_dropped_on_update():
    """A model keeps the host plans of the structural states it met (base and switched topologies): a
    second calculation reuses them and gives the same bytes. After model.update changes the base state
    (a line opened for good) the plans are rebuilt: the results equal a fresh model built with the line
    open."""
    case = synthetic.year_20k()
    sw = np.flatnonzero(np.diff(case.update["line"]["indptr"]) > 0)
    upd = _subset(case, list(range(0, 40)) + list(sw[:8]))
    sel = {"node": ["u_pu", "u_angle"], "line": ["p_from", "q_from"]}
    model = PowerGridModel(case.input)
    first = model.calculate_power_flow(update_data=upd, output_component_types=sel)
    again = model.calculate_power_flow(update_data=upd, output_component_types=sel)
    for comp in sel:
        for attr in sel[comp]:
            assert first[comp][attr].tobytes() == again[comp][attr].tobytes(), (comp, attr)
    # open one feeder line in the base state (not a line any scenario switches)
    switched = set(upd["line"]["data"]["id"].tolist())
    line = next(i for i in case.input["line"]["id"][100:] if int(i) not in switched)
    from power_grid_model_amd import initialize_array
    lu = initialize_array("update", "line", 1)
    lu["id"] = line
    lu["from_status"] = 0
    lu["to_status"] = 0
    model.update(update_data={"line": lu})
    inp = {k: v.copy() for k, v in case.input.items()}
    inp["line"]["from_status"][inp["line"]["id"] == line] = 0
    inp["line"]["to_status"][inp["line"]["id"] == line] = 0
    fresh = PowerGridModel(inp).calculate_power_flow(update_data=upd, output_component_types=sel)
    after = model.calculate_power_flow(update_data=upd, output_component_types=sel)
    for comp in sel:
        for attr in sel[comp]:
            assert after[comp][attr].tobytes() == fresh[comp][attr].tobytes(), (comp, attr)
    assert after["node"]["u_pu"].tobytes() != first["node"]["u_pu"].tobytes()
