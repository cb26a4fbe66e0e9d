"""Seeded synthetic grids and scenario batches of the benchmark workloads.

The five workloads (input + update datasets in this package's dataset format):
  radial_200      200-node radial feeder, 100 load scenarios (sym)
  meshed_1000     1000-node tree + 50 loop lines, 500 loads, 4096 scenarios with 10% generation (sym)
  asym_500        500-node radial feeder, 400 three-phase loads, 2048 scenarios (asym)
  feeders_64      64 independent feeders of 50..2000 nodes, 128 scenarios each (sym)
  year_20k        20 002-node MV grid: 110/10.5 kV transformer, 50 x 400-node feeders, 200 open ties,
                  12 000 loads, 8760 hourly scenarios, 10% of them switching 1-3 ties and 1-3 lines (sym)
  year_20k_tap    year_20k with an on-load tap changer on the 110/10.5 kV transformer (taps -9..9 of
                  1.5% at the 110 kV side) and a transformer_tap_regulator holding the 10.5 kV busbar
                  at 10.5 kV +- 1.2% (the automatic tap changer's workload)

As literally specified these grids are loaded beyond voltage collapse (the Newton-Raphson method of
the reference would report IterationDiverge for most scenarios), so every load of a workload is
multiplied by one documented scale `kappa`, recorded in the returned case.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .meta import initialize_array

U_MV = 10.5e3


@dataclass
class Case:
    name: str
    input: dict
    update: dict | None
    n_scenarios: int
    symmetric: bool = True
    kappa: float = 1.0
    seed: int = 0
    info: dict = field(default_factory=dict)


def _tree_parents(rng, n):
    # parent of node i drawn uniformly from [max(0, i-5), i-1]
    par = np.zeros(n, dtype=np.int64)
    for i in range(1, n):
        par[i] = rng.integers(max(0, i - 5), i)
    return par


def _depths(par):
    d = np.zeros(len(par), dtype=np.int64)
    for i in range(1, len(par)):
        d[i] = d[par[i]] + 1
    return d


def _line_params(rng, n, arr):
    arr["r1"] = rng.uniform(0.05, 0.30, n)
    arr["x1"] = rng.uniform(0.02, 0.15, n)
    arr["c1"] = rng.uniform(1e-8, 5e-8, n)
    arr["tan1"] = 0.0
    arr["r0"] = 3 * arr["r1"]
    arr["x0"] = 3 * arr["x1"]
    arr["c0"] = arr["c1"]
    arr["tan0"] = 0.0
    arr["i_n"] = 500.0


def _feeder(rng, n, node_offset, line_offset, n_loops=0):
    """n-node tree (node 0 is the feeder root) with n_loops extra lines between nodes of depth >= 3."""
    par = _tree_parents(rng, n)
    depth = _depths(par)
    edges = [(int(par[i]), i) for i in range(1, n)]
    existing = {tuple(sorted(e)) for e in edges}
    deep = np.flatnonzero(depth >= 3)
    loops = []
    tries = 0
    while len(loops) < n_loops and len(deep) >= 2 and tries < 100 * (n_loops + 1):
        tries += 1
        a, b = rng.choice(deep, 2, replace=False)
        key = (int(min(a, b)), int(max(a, b)))
        if key in existing:
            continue
        existing.add(key)
        loops.append(key)
    all_edges = edges + loops
    line = initialize_array("input", "line", len(all_edges))
    line["id"] = line_offset + np.arange(len(all_edges))
    line["from_node"] = node_offset + np.array([e[0] for e in all_edges])
    line["to_node"] = node_offset + np.array([e[1] for e in all_edges])
    line["from_status"] = 1
    line["to_status"] = 1
    _line_params(rng, len(all_edges), line)
    return line, depth


def _nodes(ids, u):
    node = initialize_array("input", "node", len(ids))
    node["id"] = ids
    node["u_rated"] = u
    return node


def _source(id_, node, u_ref=1.05, sk=1e10):
    src = initialize_array("input", "source", 1)
    src["id"] = id_
    src["node"] = node
    src["status"] = 1
    src["u_ref"] = u_ref
    src["sk"] = sk
    return src


def radial_200(seed: int = 0, n_scenarios: int = 100, n_node: int = 200, kappa: float = 0.1) -> Case:
    rng = np.random.default_rng(seed)
    node = _nodes(np.arange(n_node), U_MV)
    line, _ = _feeder(rng, n_node, 0, 10_000_000)
    load = initialize_array("input", "sym_load", n_node - 1)
    load["id"] = 20_000_000 + np.arange(n_node - 1)
    load["node"] = np.arange(1, n_node)
    load["status"] = 1
    load["type"] = 0
    p = rng.uniform(1e4, 1e5, n_node - 1) * kappa
    load["p_specified"] = p
    load["q_specified"] = 0.3 * p
    src = _source(30_000_000, 0)
    inp = {"node": node, "line": line, "source": src, "sym_load": load}
    scale = rng.lognormal(0.0, 0.3, n_scenarios)
    upd = initialize_array("update", "sym_load", (n_scenarios, n_node - 1))
    upd["id"] = load["id"][None, :]
    upd["status"] = 1
    upd["p_specified"] = scale[:, None] * p[None, :]
    upd["q_specified"] = scale[:, None] * 0.3 * p[None, :]
    return Case("radial_200", inp, {"sym_load": upd}, n_scenarios, True, kappa, seed)


def radial_feeder_case(n_node: int = 60, n_scenarios: int = 32, seed: int = 0, kappa: float = 0.1) -> Case:
    c = radial_200(seed=seed, n_scenarios=n_scenarios, n_node=n_node, kappa=kappa)
    c.name = f"radial_{n_node}"
    return c


def _meshed(rng, n_node, n_loops, n_load, node_offset, line_offset, load_offset, kappa):
    line, _ = _feeder(rng, n_node, node_offset, line_offset, n_loops)
    load_nodes = node_offset + 1 + rng.choice(n_node - 1, size=n_load, replace=n_load > n_node - 1)
    load = initialize_array("input", "sym_load", n_load)
    load["id"] = load_offset + np.arange(n_load)
    load["node"] = load_nodes
    load["status"] = 1
    load["type"] = 0
    p = rng.uniform(2e4, 2e5, n_load) * kappa
    qp = rng.uniform(0.1, 0.4, n_load)
    load["p_specified"] = p
    load["q_specified"] = qp * p
    return line, load, p, qp


def _meshed_scenarios(rng, load, p, qp, n_scenarios, kappa):
    n = len(load)
    scale = rng.lognormal(0.0, 0.3, (n_scenarios, n))
    ps = scale * p[None, :]
    gen = rng.random((n_scenarios, n)) < 0.10
    ps = np.where(gen, rng.uniform(-1.5e5, 0.0, (n_scenarios, n)) * kappa, ps)
    upd = initialize_array("update", "sym_load", (n_scenarios, n))
    upd["id"] = load["id"][None, :]
    upd["status"] = 1
    upd["p_specified"] = ps
    upd["q_specified"] = ps * qp[None, :]
    return upd


def meshed_1000(seed: int = 1, n_scenarios: int = 4096, kappa: float = 0.3) -> Case:
    rng = np.random.default_rng(seed)
    n_node = 1000
    node = _nodes(np.arange(n_node), U_MV)
    line, load, p, qp = _meshed(rng, n_node, 50, 500, 0, 10_000_000, 20_000_000, kappa)
    src = _source(30_000_000, 0)
    upd = _meshed_scenarios(rng, load, p, qp, n_scenarios, kappa)
    inp = {"node": node, "line": line, "source": src, "sym_load": load}
    return Case("meshed_1000", inp, {"sym_load": upd}, n_scenarios, True, kappa, seed)


def asym_500(seed: int = 2, n_scenarios: int = 2048, kappa: float = 0.025) -> Case:
    rng = np.random.default_rng(seed)
    n_node = 500
    node = _nodes(np.arange(n_node), U_MV)
    line, _ = _feeder(rng, n_node, 0, 10_000_000)
    n_load = 400
    load = initialize_array("input", "asym_load", n_load)
    load["id"] = 20_000_000 + np.arange(n_load)
    load["node"] = 1 + rng.choice(n_node - 1, size=n_load, replace=False)
    load["status"] = 1
    load["type"] = 0
    p = rng.uniform(0, 3e4, (n_load, 3)) * kappa
    load["p_specified"] = p
    load["q_specified"] = 0.25 * p
    src = _source(30_000_000, 0)
    scale = rng.lognormal(0.0, 0.4, (n_scenarios, n_load, 3))
    upd = initialize_array("update", "asym_load", (n_scenarios, n_load))
    upd["id"] = load["id"][None, :]
    upd["status"] = 1
    upd["p_specified"] = scale * p[None]
    upd["q_specified"] = 0.25 * scale * p[None]
    inp = {"node": node, "line": line, "source": src, "asym_load": load}
    return Case("asym_500", inp, {"asym_load": upd}, n_scenarios, False, kappa, seed)


def feeders_64(seed: int = 3, n_scenarios: int = 128, kappa: float = 0.06, n_feeder: int = 64) -> Case:
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(50), np.log(2000), n_feeder))).astype(int)
    nodes, lines, loads, srcs, ps, qps = [], [], [], [], [], []
    off = 0
    for f, n in enumerate(sizes):
        nodes.append(_nodes(off + np.arange(n), U_MV))
        line, load, p, qp = _meshed(rng, n, int(round(0.05 * n)), int(round(0.5 * n)), off,
                                     10_000_000 + len(np.concatenate(lines)["id"]) if lines else 10_000_000,
                                     20_000_000 + sum(len(x) for x in loads), kappa)
        lines.append(line)
        loads.append(load)
        ps.append(p)
        qps.append(qp)
        srcs.append(_source(30_000_000 + f, off))
        off += n
    load = np.concatenate(loads)
    upd = _meshed_scenarios(rng, load, np.concatenate(ps), np.concatenate(qps), n_scenarios, kappa)
    inp = {"node": np.concatenate(nodes), "line": np.concatenate(lines), "source": np.concatenate(srcs),
           "sym_load": load}
    return Case("feeders_64", inp, {"sym_load": upd}, n_scenarios, True, kappa, seed,
                info={"feeder_sizes": sizes.tolist(), "instances": int(n_feeder * n_scenarios)})


def year_20k(seed: int = 4, n_scenarios: int = 8760, kappa: float = 0.04, n_feeder: int = 50,
             feeder_nodes: int = 400, n_ties: int = 200, n_load: int = 12000, switching_fraction: float = 0.1) -> Case:
    rng = np.random.default_rng(seed)
    # node 0: 110 kV, node 1: 10.5 kV busbar, feeders follow
    n_node = 2 + n_feeder * feeder_nodes
    node = initialize_array("input", "node", n_node)
    node["id"] = np.arange(n_node)
    node["u_rated"] = U_MV
    node["u_rated"][0] = 110e3
    src = _source(30_000_000, 0)
    tr = initialize_array("input", "transformer", 1)
    tr["id"] = 40_000_000
    tr["from_node"] = 0
    tr["to_node"] = 1
    tr["from_status"] = 1
    tr["to_status"] = 1
    tr["u1"] = 110e3
    tr["u2"] = U_MV
    tr["sn"] = 40e6
    tr["uk"] = 0.12
    tr["pk"] = 1e5
    tr["i0"] = 0.0
    tr["p0"] = 0.0
    tr["winding_from"] = 1
    tr["winding_to"] = 1
    tr["clock"] = 0
    tr["tap_side"] = 0
    tr["tap_pos"] = 0
    tr["tap_min"] = 0
    tr["tap_max"] = 0
    tr["tap_nom"] = 0
    tr["tap_size"] = 0.0
    lines = []
    feeder_of = np.full(n_node, -1)
    lid = 10_000_000
    for f in range(n_feeder):
        off = 2 + f * feeder_nodes
        feeder_of[off:off + feeder_nodes] = f
        line, _ = _feeder(rng, feeder_nodes, off, lid)
        lid += len(line)
        lines.append(line)
        root = initialize_array("input", "line", 1)
        root["id"] = lid
        lid += 1
        root["from_node"] = 1
        root["to_node"] = off
        root["from_status"] = 1
        root["to_status"] = 1
        _line_params(rng, 1, root)
        lines.append(root)
    feeder_lines = np.concatenate(lines)
    # normally open ties between nodes of different feeders
    ties = initialize_array("input", "line", n_ties)
    ties["id"] = lid + np.arange(n_ties)
    a = 2 + rng.integers(0, n_feeder * feeder_nodes, n_ties)
    b = 2 + rng.integers(0, n_feeder * feeder_nodes, n_ties)
    same = feeder_of[a] == feeder_of[b]
    b[same] = 2 + (b[same] - 2 + feeder_nodes) % (n_feeder * feeder_nodes)
    ties["from_node"] = a
    ties["to_node"] = b
    ties["from_status"] = 0
    ties["to_status"] = 0
    _line_params(rng, n_ties, ties)
    line = np.concatenate([feeder_lines, ties])
    load = initialize_array("input", "sym_load", n_load)
    load["id"] = 20_000_000 + np.arange(n_load)
    load["node"] = 2 + rng.integers(0, n_feeder * feeder_nodes, n_load)
    load["status"] = 1
    load["type"] = 0
    p = rng.uniform(1e4, 1e5, n_load) * kappa
    load["p_specified"] = p
    load["q_specified"] = 0.3 * p
    h = np.arange(n_scenarios)
    daily = 1.0 + 0.4 * np.sin(2 * np.pi * h / 24.0)
    scale = daily[:, None] * rng.lognormal(0.0, 0.2, (n_scenarios, n_load))
    upd_load = initialize_array("update", "sym_load", (n_scenarios, n_load))
    upd_load["id"] = load["id"][None, :]
    upd_load["status"] = 1
    upd_load["p_specified"] = scale * p[None, :]
    upd_load["q_specified"] = scale * 0.3 * p[None, :]
    # switching scenarios: close 1-3 ties and open 1-3 feeder lines
    switching = np.flatnonzero(rng.random(n_scenarios) < switching_fraction)
    indptr = np.zeros(n_scenarios + 1, dtype=np.int64)
    rows = []
    counts = np.zeros(n_scenarios, dtype=np.int64)
    n_feeder_lines = len(feeder_lines)
    for s in switching:
        close = rng.choice(n_ties, rng.integers(1, 4), replace=False)
        open_ = rng.choice(n_feeder_lines, rng.integers(1, 4), replace=False)
        r = initialize_array("update", "line", len(close) + len(open_))
        r["id"] = np.concatenate([ties["id"][close], feeder_lines["id"][open_]])
        r["from_status"] = np.concatenate([np.ones(len(close)), np.zeros(len(open_))])
        r["to_status"] = r["from_status"]
        rows.append(r)
        counts[s] = len(r)
    indptr[1:] = np.cumsum(counts)
    line_data = np.concatenate(rows) if rows else initialize_array("update", "line", 0)
    # keep rows in scenario order
    order = []
    pos = 0
    by_scn = {}
    for s, r in zip(switching, rows):
        by_scn[int(s)] = r
    line_data = np.concatenate([by_scn[s] for s in sorted(by_scn)]) if by_scn else line_data
    del order, pos
    inp = {"node": node, "line": line, "transformer": tr, "source": src, "sym_load": load}
    upd = {"sym_load": upd_load, "line": {"indptr": indptr, "data": line_data}}
    return Case("year_20k", inp, upd, n_scenarios, True, kappa, seed,
                info={"switching_scenarios": int(len(switching))})


def year_20k_tap(seed: int = 4, n_scenarios: int = 8760, kappa: float = 0.04, u_band: float = 252.0,
                 **kwargs) -> Case:
    case = year_20k(seed=seed, n_scenarios=n_scenarios, kappa=kappa, **kwargs)
    tr = case.input["transformer"]
    tr["tap_min"] = -9
    tr["tap_max"] = 9
    tr["tap_nom"] = 0
    tr["tap_pos"] = 0
    tr["tap_size"] = 0.015 * 110e3
    reg = initialize_array("input", "transformer_tap_regulator", 1)
    reg["id"] = 50_000_000
    reg["regulated_object"] = tr["id"][0]
    reg["status"] = 1
    reg["control_side"] = 1
    reg["u_set"] = U_MV
    reg["u_band"] = u_band
    reg["line_drop_compensation_r"] = 0.0
    reg["line_drop_compensation_x"] = 0.0
    case.input["transformer_tap_regulator"] = reg
    case.name = "year_20k_tap"
    return case


CONFIGS = {
    "radial_200": radial_200,
    "meshed_1000": meshed_1000,
    "asym_500": asym_500,
    "feeders_64": feeders_64,
    "year_20k": year_20k,
    "year_20k_tap": year_20k_tap,
}


def state_estimation_case(case: Case, pf_output: dict, seed: int = 7, n_scenarios: int | None = None,
                          voltage_share: float = 0.05, flow_share: float = 0.1, noise: float = 1e-3) -> Case:
    """A state-estimation workload on a power-flow case: sensors placed from its structure and measurements taken
    from its power-flow results (pf_output, a batch over the case's scenarios) with relative Gaussian noise.

    Sensors: a voltage phasor sensor at the source node, magnitude-only voltage sensors at a share of the other
    nodes, a power sensor on every load and on the source (every injection measured, so the grid is observable),
    and from-side power sensors on a share of the lines. One scenario per power-flow scenario (the first
    n_scenarios): the update batch carries every sensor's measured values."""
    rng = np.random.default_rng(seed)
    inp = {k: v.copy() for k, v in case.input.items()}
    node = inp["node"]
    n_node = len(node)
    src = inp["source"]
    load = inp["sym_load"]
    line = inp["line"]
    s_total = pf_output["node"].shape[0] if n_scenarios is None else n_scenarios
    src_node = int(src["node"][0])
    others = np.setdiff1d(node["id"], [src_node])
    vnodes = np.concatenate([[src_node], rng.choice(others, size=max(1, int(voltage_share * len(others))), replace=False)])
    flines = rng.choice(len(line), size=max(1, int(flow_share * len(line))), replace=False)
    next_id = int(max(int(node["id"].max()), int(line["id"].max()), int(load["id"].max()), int(src["id"].max()))) + 1
    vs = initialize_array("input", "sym_voltage_sensor", len(vnodes))
    vs["id"] = next_id + np.arange(len(vnodes))
    vs["measured_object"] = vnodes
    vs["u_sigma"] = 0.001 * U_MV
    next_id += len(vnodes)
    n_ps = len(load) + 1 + len(flines)
    ps = initialize_array("input", "sym_power_sensor", n_ps)
    ps["id"] = next_id + np.arange(n_ps)
    ps["measured_object"][: len(load)] = load["id"]
    ps["measured_terminal_type"][: len(load)] = 4  # load
    ps["measured_object"][len(load)] = src["id"][0]
    ps["measured_terminal_type"][len(load)] = 2  # source
    ps["measured_object"][len(load) + 1:] = line["id"][flines]
    ps["measured_terminal_type"][len(load) + 1:] = 0  # branch from side
    ps["power_sigma"] = 1e3
    # measurements of every scenario from the power-flow results
    node_pos = {int(i): k for k, i in enumerate(pf_output["node"]["id"][0])}
    vpos = np.array([node_pos[int(i)] for i in vnodes])
    u = pf_output["node"]["u"][:s_total][:, vpos]
    ang = pf_output["node"]["u_angle"][:s_total][:, vpos]
    vs_upd = initialize_array("update", "sym_voltage_sensor", (s_total, len(vnodes)))
    vs_upd["id"] = vs["id"][None, :]
    vs_upd["u_measured"] = u * (1.0 + noise * rng.standard_normal(u.shape))
    vs_upd["u_angle_measured"][:, 0] = ang[:, 0]
    p = np.concatenate([pf_output["sym_load"]["p"][:s_total], pf_output["source"]["p"][:s_total],
                        pf_output["line"]["p_from"][:s_total][:, flines]], axis=1)
    q = np.concatenate([pf_output["sym_load"]["q"][:s_total], pf_output["source"]["q"][:s_total],
                        pf_output["line"]["q_from"][:s_total][:, flines]], axis=1)
    ps_upd = initialize_array("update", "sym_power_sensor", (s_total, n_ps))
    ps_upd["id"] = ps["id"][None, :]
    ps_upd["p_measured"] = p + noise * np.abs(p).mean() * rng.standard_normal(p.shape)
    ps_upd["q_measured"] = q + noise * np.abs(q).mean() * rng.standard_normal(q.shape)
    inp["sym_voltage_sensor"] = vs
    inp["sym_power_sensor"] = ps
    upd = {"sym_voltage_sensor": vs_upd, "sym_power_sensor": ps_upd}
    return Case("se_" + case.name, inp, upd, s_total, True, case.kappa, seed)
