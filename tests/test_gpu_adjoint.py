"""Adjoint gradients of the HIP path against finite differences of the same solver.

Objective per scenario: L = sum over energised nodes (u_pu - 1)^2 + sum over lines (p_from + p_to) in MW.
Gradients by load p_specified / q_specified and by line r1 / x1 come from one transposed solve with the
Newton-Raphson Jacobian factorised at the converged state.

The finite differences: every side is solved to 1e-12 pu in FP64. Even so the converged states of a
200-bus feeder carry rounding of about 1e-11 in L (the CPU oracle and the GPU disagree at that level),
so a central difference at h = 1e-6 relative has a few percent of noise for these gradients. The test
therefore uses h = 1e-2 relative with Richardson extrapolation, D = (8 D1(h) - D1(2h)) / 6 over the
one-sided central differences D1: truncation O(h^4) ~ 1e-8, noise ~ 1e-6 of a typical gradient. The bar
is 2e-6 relative, with an absolute floor of 1e-4 times the largest gradient of the kind (gradients that
small are below the differences' noise).
"""

import numpy as np
import pytest

pgm = pytest.importorskip("power_grid_model_amd")
from power_grid_model_amd import DeviceBatch, PowerGridModel, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

OUTPUT = {"node": ["u_pu", "energized"], "line": ["p_from", "p_to"]}


def _objective(model, update, precision="fp64", tol=1e-12):
    batch = DeviceBatch(model, update, error_tolerance=tol, max_iterations=30, factor_precision=precision)
    batch.set_output(OUTPUT)
    batch.run()
    out = batch.fetch()
    u, en = out["node"]["u_pu"], out["node"]["energized"]
    loss = np.sum(np.where(en == 1, (u - 1.0) ** 2, 0.0), axis=-1)
    loss = loss + np.sum(out["line"]["p_from"] + out["line"]["p_to"], axis=-1) / 1e6
    return loss, batch


def _case():
    """radial_200 with a third of the loads const_z and a third const_i (types 1 and 2)."""
    case = synthetic.radial_200(n_scenarios=40)
    inp = {k: v.copy() for k, v in case.input.items()}
    n = len(inp["sym_load"])
    inp["sym_load"]["type"] = np.arange(n) % 3
    return inp, case.update


H_REL = 1e-2


def _richardson(lp, lm, lp2, lm2, h):
    """(8 (L(+h) - L(-h)) - (L(+2h) - L(-2h))) / (12 h)"""
    return (8.0 * (lp - lm) - (lp2 - lm2)) / (12.0 * h)


def _close(adj, fd, name):
    floor = 1e-4 * max(np.max(np.abs(adj)), np.max(np.abs(fd)))
    np.testing.assert_allclose(adj, fd, rtol=2e-6, atol=floor, err_msg=name)


def test_adjoint_load_gradients_match_finite_differences():
    inp, upd = _case()
    model = PowerGridModel(inp)
    loss, batch = _objective(model, upd)
    grads = batch.adjoint()["sym_load"]
    rng = np.random.default_rng(1)
    scn = 7
    loads = rng.choice(len(inp["sym_load"]), 12, replace=False)
    base = upd["sym_load"][scn]
    for attr in ("p_specified", "q_specified"):
        rows = []
        steps = []
        for l in loads:
            h = H_REL * abs(base[attr][l])
            steps.append(h)
            for mult in (1.0, -1.0, 2.0, -2.0):
                r = base.copy()
                r[attr][l] += mult * h
                rows.append(r)
        pert, _ = _objective(model, {"sym_load": np.stack(rows)})
        fd = _richardson(pert[0::4], pert[1::4], pert[2::4], pert[3::4], np.array(steps))
        _close(grads[attr][scn, loads], fd, attr)


def test_adjoint_line_gradients_match_finite_differences():
    inp, upd = _case()
    model = PowerGridModel(inp)
    _, batch = _objective(model, upd)
    grads = batch.adjoint()["line"]
    rng = np.random.default_rng(2)
    scn = 11
    lines = rng.choice(len(inp["line"]), 8, replace=False)
    one = {"sym_load": upd["sym_load"][scn:scn + 1]}
    for attr, key in (("r1", "r1"), ("x1", "x1")):
        fd = []
        for b in lines:
            h = H_REL * abs(inp["line"][attr][b])
            vals = []
            for mult in (1.0, -1.0, 2.0, -2.0):
                mod = {k: v.copy() for k, v in inp.items()}
                mod["line"][attr][b] += mult * h
                val, _ = _objective(PowerGridModel(mod), one)
                vals.append(val[0])
            fd.append(_richardson(*vals, h))
        _close(grads[key][scn, lines], np.array(fd), attr)


def test_adjoint_mixed_precision_against_fp64():
    """The adjoint factorises in FP64 also for a mixed-precision batch: its gradients equal the FP64
    batch's up to the two converged states' difference (both within 1e-8 pu)."""
    inp, upd = _case()
    model = PowerGridModel(inp)
    _, b64 = _objective(model, upd, "fp64", 1e-8)
    _, b32 = _objective(model, upd, "fp32", 1e-8)
    g64, g32 = b64.adjoint(), b32.adjoint()
    for comp, attrs in (("sym_load", ("p_specified", "q_specified")), ("line", ("r1", "x1"))):
        for a in attrs:
            ref = g64[comp][a]
            np.testing.assert_allclose(g32[comp][a], ref, rtol=1e-5, atol=1e-6 * np.max(np.abs(ref)),
                                       err_msg=f"{comp}.{a}")


def test_adjoint_rejects_unsupported_batches():
    inp, upd = _case()
    model = PowerGridModel(inp)
    batch = DeviceBatch(model, {"sym_load": upd["sym_load"][:8]})  # 8 scenarios: a packed group
    batch.run()
    with pytest.raises(Exception, match="more than 32 scenarios"):
        batch.adjoint()


def test_torch_function_directional_derivative():
    """PowerFlowLoss.backward (the adjoint) against a Richardson-extrapolated directional difference of the
    weighted total loss along a random direction in every load p and every line r1 and x1."""
    torch = pytest.importorskip("torch")
    from power_grid_model_amd.autograd import PowerFlowLoss

    inp, upd = _case()
    loads = upd["sym_load"]
    p0 = torch.tensor(loads["p_specified"], dtype=torch.float64)
    q0 = torch.tensor(loads["q_specified"], dtype=torch.float64)
    r0 = torch.tensor(inp["line"]["r1"], dtype=torch.float64)
    x0 = torch.tensor(inp["line"]["x1"], dtype=torch.float64)
    p, q, r, x = (t.clone().requires_grad_(True) for t in (p0, q0, r0, x0))
    w = torch.linspace(0.5, 1.5, p0.shape[0], dtype=torch.float64)
    loss = PowerFlowLoss.apply(p, q, r, x, inp, "fp32", 1e-8, 20)
    (loss * w).sum().backward()
    rng = np.random.default_rng(5)
    dp = torch.tensor(rng.standard_normal(p0.shape), dtype=torch.float64) * p0 * 1e-2
    dr = torch.tensor(rng.standard_normal(r0.shape), dtype=torch.float64) * r0 * 1e-2
    dx = torch.tensor(rng.standard_normal(x0.shape), dtype=torch.float64) * x0 * 1e-2

    def total(t):
        with torch.no_grad():
            val = PowerFlowLoss.apply(p0 + t * dp, q0, r0 + t * dr, x0 + t * dx, inp, "fp64", 1e-12, 30)
        return float((val * w).sum())

    fd = _richardson(total(1.0), total(-1.0), total(2.0), total(-2.0), 1.0)
    ad = float((p.grad * dp).sum() + (r.grad * dr).sum() + (x.grad * dx).sum())
    assert abs(ad - fd) <= 2e-6 * abs(fd), (ad, fd)
    assert q.grad.shape == q0.shape and torch.isfinite(q.grad).all()


def test_adjoint_on_meshed_grid_with_fill_ins():
    """meshed_1000 (loops, fill-ins in the factors): load and line gradients against finite differences."""
    case = synthetic.meshed_1000(n_scenarios=40)
    inp, upd = case.input, case.update
    model = PowerGridModel(inp)
    _, batch = _objective(model, upd)
    grads = batch.adjoint()
    rng = np.random.default_rng(4)
    scn = 3
    loads = rng.choice(len(inp["sym_load"]), 6, replace=False)
    base = upd["sym_load"][scn]
    rows, steps = [], []
    for l in loads:
        h = H_REL * abs(base["p_specified"][l])
        steps.append(h)
        for mult in (1.0, -1.0, 2.0, -2.0):
            r = base.copy()
            r["p_specified"][l] += mult * h
            rows.append(r)
    pert, _ = _objective(model, {"sym_load": np.stack(rows)})
    fd = _richardson(pert[0::4], pert[1::4], pert[2::4], pert[3::4], np.array(steps))
    _close(grads["sym_load"]["p_specified"][scn, loads], fd, "p_specified")
    one = {"sym_load": upd["sym_load"][scn:scn + 1]}
    lines = rng.choice(len(inp["line"]), 4, replace=False)
    fd = []
    for b in lines:
        h = H_REL * abs(inp["line"]["x1"][b])
        vals = []
        for mult in (1.0, -1.0, 2.0, -2.0):
            mod = {k: v.copy() for k, v in inp.items()}
            mod["line"]["x1"][b] += mult * h
            val, _ = _objective(PowerGridModel(mod), one)
            vals.append(val[0])
        fd.append(_richardson(*vals, h))
    _close(grads["line"]["x1"][scn, lines], np.array(fd), "x1")


# ------------------------------------------------------------------------------------------------
# the uneven-feeder workload at full width: 64 feeders of 50-2000 nodes, gradients against the oracle
# ------------------------------------------------------------------------------------------------
FEEDER_OUTPUT = {"node": ["u_pu", "u_angle", "energized"], "line": ["p_from", "p_to"]}


def _feeder_maps(case):
    sizes = case.info["feeder_sizes"]
    feeder_of_node = np.repeat(np.arange(len(sizes)), sizes)
    inp = case.input
    return feeder_of_node, feeder_of_node[inp["line"]["from_node"]], feeder_of_node[inp["sym_load"]["node"]]


def _feeder_losses(node_out, line_tab, feeder_of_node, feeder_of_line, n_feeder):
    """Per scenario and feeder, L_f = sum over its energised nodes (u_pu - 1)^2 + its line losses in MW.
    A line's p_from + p_to equals g_s |u_from - u_to|^2 when its shunt is lossless (tan1 = 0 here); the
    losses are formed that way, from the node voltages, because the sum of the two large end flows loses
    about 1e-13 MW per line to cancellation, which swamps the finite differences."""
    u, en = node_out["u_pu"], node_out["energized"]
    uc = u * np.exp(1j * node_out["u_angle"])
    z_base = (10.5e3 ** 2) / 1e6
    g = (1.0 / ((line_tab["r1"] + 1j * line_tab["x1"]) / z_base)).real
    loss = g * np.abs(uc[:, line_tab["from_node"]] - uc[:, line_tab["to_node"]]) ** 2
    nodes = np.where(en == 1, (u - 1.0) ** 2, 0.0)
    return np.stack([np.bincount(feeder_of_node, weights=nodes[s], minlength=n_feeder) +
                     np.bincount(feeder_of_line, weights=loss[s], minlength=n_feeder) for s in range(u.shape[0])])


def _pick(owner, n_feeder, k):
    """The first k elements of every feeder: array (k, n_feeder) of element indices."""
    return np.stack([np.flatnonzero(owner == f)[:k] for f in range(n_feeder)], axis=1)


FD_STEPS = (1e-1, 5e-2, 2.5e-2)
MULTS = (1.0, -1.0, 2.0, -2.0)


def _fd_from(vals, steps):
    """Richardson-extrapolated central differences from L at (+h, -h, +2h, -2h)."""
    return _richardson(vals[0], vals[1], vals[2], vals[3], steps)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_feeders_64_against_oracle_finite_differences(precision):
    """feeders_64 at full width (64 feeders, 31.8k buses, 40 scenarios, solved to 1e-8 pu): the adjoint's
    gradients by 8 load p, 8 load q, 8 line r1 and 8 line x1 of every feeder (32 per feeder, 2048 in all) in
    scenario 2, against finite differences of the CPU oracle solved to 1e-12 pu in FP64 (an FP64 reference
    gradient independent of the GPU solver), and the load gradients also against finite differences of the
    GPU solver.

    Finite differences: a parameter of one feeder changes that feeder's L_f only (the feeders are separate
    math models), so the j-th parameter of every feeder is perturbed in the same solve. L_f is known to about
    1e-12 at best, so central differences at h = 1e-6 relative carry up to 1e-4 relative noise on these
    gradients; each reference gradient is instead a Richardson extrapolation (O(h^4)) at h = 1e-1 relative.
    Its error is estimated per parameter from the same extrapolation at h = 5e-2 and 2.5e-2: truncation
    shrinks with h and rounding noise grows as 1/h, so the larger of the two successive differences bounds
    both (one difference alone can be small by chance, when the noise at two steps happens to agree).
    Bar: |adjoint - FD| <= 1e-6 |FD| + 2 x that estimate, plus for line gradients the adjoint's own rounding
    floor: a line's gradient is formed from the adjoint variables at its two ends, whose difference across a
    line that carries little current is far smaller than the variables themselves, so their FP64 rounding
    (the solve of a feeder's whole system) leaves an absolute error of a few 1e-9 of the feeder's largest
    line gradient (measured: 2.3e-9 on the worst of the 512 sampled lines, one whose gradient is 1/1000 of
    its feeder's largest); the floor is 5e-9 of the largest |FD| among the feeder's sampled lines of that
    kind. The line gradients span eight orders of magnitude
    (lines at feeder ends carry almost nothing), so the estimate resolves 1e-6 relative for most load and
    many line gradients, not all; the test asserts the resolved share it reaches on this workload."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle_lib as ol

    case = synthetic.feeders_64(n_scenarios=40)
    inp, upd = case.input, case.update
    n_feeder = len(case.info["feeder_sizes"])
    fon, fol, fold = _feeder_maps(case)
    model = PowerGridModel(inp)
    batch = DeviceBatch(model, upd, factor_precision=precision)
    batch.set_output({"node": ["u_pu", "energized"], "line": ["p_from", "p_to"]})
    batch.run()
    grads = batch.adjoint()
    scn = 2
    base = upd["sym_load"][scn]
    loads, lines = _pick(fold, n_feeder, 8), _pick(fol, n_feeder, 8)

    def oracle_losses(inp_mod, rows):
        out, it, st, _ = ol.power_flow(inp_mod, {"sym_load": np.stack(rows)}, len(rows), err_tol=1e-12,
                                       max_iter=30, threading=0)
        assert (st == 0).all()
        return _feeder_losses(out["node"], inp_mod["line"], fon, fol, n_feeder)

    def gpu_losses(inp_mod, rows):
        b = DeviceBatch(PowerGridModel(inp_mod), {"sym_load": np.stack(rows)}, error_tolerance=1e-12,
                        max_iterations=30)
        b.set_output(FEEDER_OUTPUT)
        b.run()
        assert (b.iterations() > 0).all()
        return _feeder_losses(b.fetch()["node"], inp_mod["line"], fon, fol, n_feeder)

    def fd_of(L, steps_abs):
        L = L.reshape(8, len(FD_STEPS), 4, n_feeder)
        return np.stack([[_fd_from(L[j, k], h * steps_abs[j]) for k, h in enumerate(FD_STEPS)] for j in range(8)])

    results = {}
    for attr in ("p_specified", "q_specified"):
        rows = []
        for j in range(8):
            for h_rel in FD_STEPS:
                for mult in MULTS:
                    r = base.copy()
                    r[attr][loads[j]] += mult * h_rel * np.abs(base[attr][loads[j]])
                    rows.append(r)
        steps = np.abs(base[attr][loads])
        results[(attr, "oracle")] = fd_of(oracle_losses(inp, rows), steps)
        results[(attr, "gpu")] = fd_of(gpu_losses(inp, rows), steps)
        results[(attr, "adjoint")] = grads["sym_load"][attr][scn][loads]
    for attr in ("r1", "x1"):
        mods = []
        for j in range(8):
            for h_rel in FD_STEPS:
                for mult in MULTS:
                    mod = {c: v.copy() for c, v in inp.items()}
                    mod["line"][attr][lines[j]] += mult * h_rel * np.abs(inp["line"][attr][lines[j]])
                    mods.append(mod)
        with ThreadPoolExecutor(8) as ex:
            lo = np.stack(list(ex.map(lambda m: oracle_losses(m, [base])[0], mods)))
        results[(attr, "oracle")] = fd_of(lo, np.abs(inp["line"][attr][lines]))
        results[(attr, "adjoint")] = grads["line"][attr][scn][lines]

    min_resolved = {"p_specified": 0.9, "q_specified": 0.85, "r1": 0.4, "x1": 0.25}
    for attr in ("p_specified", "q_specified", "r1", "x1"):
        adj = results[(attr, "adjoint")]
        for who in ("oracle", "gpu"):
            if (attr, who) not in results:
                continue
            fd = results[(attr, who)]
            est = np.maximum(np.abs(fd[:, 0] - fd[:, 1]), np.abs(fd[:, 1] - fd[:, 2]))
            resolved = est <= 1e-6 * np.abs(fd[:, 0])
            floor = 5e-9 * np.max(np.abs(fd[:, 0]), axis=0, keepdims=True) if attr in ("r1", "x1") else 0.0
            err = np.abs(adj - fd[:, 0])
            rel = err / np.abs(fd[:, 0])
            print(f"{precision} {attr} vs {who} FD: resolved at 1e-6 {np.mean(resolved):.0%}, max |adj - fd| / |fd| "
                  f"there {np.max(rel[resolved]):.1e}; largest |adj - fd| / (1e-6 |fd| + 2 est) "
                  f"{np.max(err / (1e-6 * np.abs(fd[:, 0]) + 2 * est + floor)):.2f}")
            assert np.mean(resolved) >= min_resolved[attr], (attr, who, np.mean(resolved))
            bad = err > 1e-6 * np.abs(fd[:, 0]) + 2 * est + floor
            assert not bad.any(), (attr, who, np.argwhere(bad)[:5], rel[bad][:5])


# ------------------------------------------------------------------------------------------------
# the adjoint layout: batches with few-scenario topologies and switching scenarios
# ------------------------------------------------------------------------------------------------
def test_adjoint_layout_small_groups_equal_plain_groups():
    """8 scenarios alone form a packed group, which the adjoint does not take; built with the adjoint layout
    they form a plain group, and their gradients equal those of the same scenarios inside a 40-scenario
    plain group."""
    inp, upd = _case()
    model = PowerGridModel(inp)
    _, big = _objective(model, upd, tol=1e-8)
    g_big = big.adjoint()
    small = DeviceBatch(model, {"sym_load": upd["sym_load"][:8]}, adjoint_layout=True)
    small.set_output(OUTPUT)
    small.run()
    g_small = small.adjoint()
    for comp, attrs in (("sym_load", ("p_specified", "q_specified")), ("line", ("r1", "x1"))):
        for a in attrs:
            ref = g_big[comp][a][:8]
            np.testing.assert_allclose(g_small[comp][a], ref, rtol=1e-12, atol=1e-14 * np.max(np.abs(ref)),
                                       err_msg=f"{comp}.{a}")


def _year_subset(case, scenarios):
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


def _year_losses(model, upd):
    b = DeviceBatch(model, upd, error_tolerance=1e-11, max_iterations=50, adjoint_layout=True)
    b.set_output(OUTPUT)
    b.run()
    assert (b.iterations() > 0).all()
    out = b.fetch()
    u, en = out["node"]["u_pu"], out["node"]["energized"]
    loss = np.sum(np.where(en == 1, (u - 1.0) ** 2, 0.0), axis=-1)
    return loss + np.sum(out["line"]["p_from"] + out["line"]["p_to"], axis=-1) / 1e6


def test_adjoint_on_switching_scenarios():
    """The year-long workload with switching scenarios (ties closed, feeder lines opened, islands) built with
    the adjoint layout: the gradients by load p and q of two switching scenarios and one plain one, against
    Richardson-extrapolated finite differences of the GPU solver (each side solved to 1e-11 pu in the same
    switching state; 20k buses carry about 1e-12 pu of rounding in their iterates). Bar as for the uneven feeders: |adjoint - FD| <= 1e-6 |FD| + 2 x the larger difference
    between the extrapolations at h, h/2 and h/4 (h = 1e-1 relative)."""
    from power_grid_model_amd import synthetic

    case = synthetic.year_20k()
    sw = np.flatnonzero(np.diff(case.update["line"]["indptr"]) > 0)
    scen = list(range(0, 16)) + list(sw[:8])
    upd = _year_subset(case, scen)
    model = PowerGridModel(case.input)
    batch = DeviceBatch(model, upd, adjoint_layout=True)
    batch.set_output(OUTPUT)
    batch.run()
    assert (batch.iterations() > 0).all()
    grads = batch.adjoint()["sym_load"]
    out = batch.fetch()
    rng = np.random.default_rng(9)
    for k in (16, 19, 3):  # two switching scenarios, one plain
        base = upd["sym_load"][k]
        energised = out["node"]["energized"][k]
        node_of = {int(i): j for j, i in enumerate(case.input["node"]["id"])}
        live = np.flatnonzero([energised[node_of[int(nd)]] == 1 for nd in case.input["sym_load"]["node"]])
        loads = rng.choice(live, 6, replace=False)
        one = _year_subset(case, [scen[k]])
        for attr in ("p_specified", "q_specified"):
            rows = []
            for l in loads:
                for h_rel in FD_STEPS:
                    for mult in MULTS:
                        r = base.copy()
                        r[attr][l] += mult * h_rel * abs(base[attr][l])
                        rows.append(r)
            n = len(rows)
            pert = {"sym_load": np.stack(rows),
                    "line": {"indptr": np.arange(n + 1, dtype=np.int64) * len(one["line"]["data"]),
                             "data": np.tile(one["line"]["data"], n)}}
            L = _year_losses(model, pert).reshape(len(loads), len(FD_STEPS), 4)
            steps = np.abs(base[attr][loads])
            fd = np.stack([[_fd_from(L[j, i], h * steps[j]) for i, h in enumerate(FD_STEPS)]
                           for j in range(len(loads))])
            est = np.maximum(np.abs(fd[:, 0] - fd[:, 1]), np.abs(fd[:, 1] - fd[:, 2]))
            adj = grads[attr][k][loads]
            err = np.abs(adj - fd[:, 0])
            print(f"scenario {scen[k]} {attr}: max |adj - fd| / |fd| {np.max(err / np.abs(fd[:, 0])):.1e}, "
                  f"est / |fd| {np.max(est / np.abs(fd[:, 0])):.1e}")
            assert (err <= 1e-6 * np.abs(fd[:, 0]) + 2 * est).all(), (scen[k], attr, err / np.abs(fd[:, 0]))
