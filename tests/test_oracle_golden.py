"""Pin the CPU oracle against the reference's own known answers.

Sources of the vectors (reference tests, restated here as data):
  * minimum degree ordering      tests/cpp_unit_tests/test_sparse_ordering.cpp:21-35
  * cycle reorder 9 / 7 nodes    tests/cpp_unit_tests/test_topology.cpp:380-469
  * Y-bus LU structure           tests/cpp_unit_tests/math_solver/test_y_bus.cpp:383-425
  * block sparse LU              tests/cpp_unit_tests/math_solver/test_sparse_lu_solver.cpp:70-250, 335-412
  * closed-form 3-bus NR         tests/cpp_unit_tests/math_solver/test_math_solver_common.hpp:62-262,
                                 test_math_solver_pf.hpp:36-153 (run at tol 1e-12)
  * validation cases             tests/data/power_flow/** (copied under tests/golden/power_flow)
  * NR iteration counts          test_math_solver_pf.hpp:118-130 (single iteration at err_tol = inf),
                                 iterative_pf_solver.hpp:53-58 (the max_iter boundary), and a dense numpy
                                 Newton brute force (tests/dense_newton.py) on 100 random grids
Each check is chosen so a dropped term / wrong sign / transposed operand fails it.
"""

import cmath
import math

import numpy as np
import pytest

import oracle_lib as ol
from dense_newton import dense_nr, random_grid
from dense_newton import from_input as dense_from_input
from pgm_formats import SYM_OUTPUT
from validation_utils import compare_result, list_cases, load_case


# ------------------------------------------------------------------------------------------------
# integer work: ordering, topology, Y-bus structure
# ------------------------------------------------------------------------------------------------
def test_minimum_degree_ordering_reference_vector():
    graph = {0: [3, 5], 1: [4, 5, 8], 2: [4, 5, 6], 3: [6, 7], 4: [6, 8], 6: [7, 8, 9], 7: [8, 9], 8: [9]}
    alpha, fills = ol.minimum_degree_ordering(graph)
    assert alpha == list(range(10))
    assert fills == [(3, 5), (4, 5), (5, 8), (5, 6), (5, 7)]


def test_cycle_reorder_9_nodes():
    br = [(0, 3), (0, 5), (1, 4), (1, 5), (1, 8), (2, 4), (2, 5), (2, 6), (3, 6), (3, 7), (4, 6), (4, 8), (6, 7),
          (6, 8), (6, 9), (7, 8), (7, 9), (8, 9)]
    group, pos, nmm, fills = ol.build_topology(10, br, [0])
    assert nmm == 1
    assert group == [0] * 10
    assert pos == list(range(10))
    assert [f[1:] for f in fills] == [(3, 5), (4, 5), (5, 8), (5, 6), (5, 7)]


def test_cycle_reorder_7_nodes():
    br = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5), (1, 4), (6, 0), (6, 2), (5, 1), (3, 1), (6, 1), (2, 1)]
    group, pos, nmm, fills = ol.build_topology(7, br, [0])
    assert group == [0] * 7
    assert pos == [0, 3, 1, 2, 4, 5, 6]
    assert [f[1:] for f in fills] == [(5, 6), (2, 6), (4, 6)]


def test_radial_order_is_reversed_dfs():
    # 0 - 1 - 2, 1 - 3 : DFS from 0 visits 0,1,2,3 -> reversed order 3,2,1,0
    group, pos, nmm, fills = ol.build_topology(4, [(0, 1), (1, 2), (1, 3)], [0])
    assert pos == [3, 2, 1, 0]
    assert fills == []


def test_topology_disconnected_and_open_branch():
    # branch 1-2 open at one side: node 2 not energised; second source on node 3 makes its own model
    group, pos, nmm, fills = ol.build_topology(4, [(0, 1), (1, 2)], [0, 3], statuses=[(1, 1), (1, 0)])
    assert nmm == 2
    assert group == [0, 0, -1, 1]
    assert pos[2] == -1 and pos[3] == 0


def test_ybus_fill_in_structure():
    s = ol.ybus_structure(3, [(1, 0), (0, 2)], [], [(1, 2)])
    assert s["row_indptr"] == [0, 3, 5, 7]
    assert s["col_indices"] == [0, 1, 2, 0, 1, 0, 2]
    assert s["bus_entry"] == [0, 4, 6]
    assert s["lu_transpose_entry"] == [0, 3, 6, 1, 4, 7, 2, 5, 8]
    assert s["y_bus_entry_indptr"] == [0, 2, 3, 4, 5, 6, 7, 8]
    assert s["row_indptr_lu"] == [0, 3, 6, 9]
    assert s["col_indices_lu"] == [0, 1, 2, 0, 1, 2, 0, 1, 2]
    assert s["diag_lu"] == [0, 4, 8]
    assert s["map_lu_y_bus"] == [0, 1, 2, 3, 4, -1, 5, -1, 6]


# ------------------------------------------------------------------------------------------------
# block sparse LU known answers
# ------------------------------------------------------------------------------------------------
FULL3 = dict(row_indptr=[0, 3, 6, 9], col_indices=[0, 1, 2, 0, 1, 2, 0, 1, 2], diag_lu=[0, 4, 8])
BLOCKS_FILL = [[[0, 1], [100, 0]], [[1, 2], [7, -1]], [[3, 4], [5, 6]], [[1, 2], [-3, 4]], [[0, 200], [3, 1]],
               [[0, 0], [0, 0]], [[5, 6], [-7, 8]], [[0, 0], [0, 0]], [[1, 0], [0, 100]]]


def test_scalar_sparse_lu():
    rc, x = ol.block_lu_solve(1, blocks=[[[v]] for v in [4, 1, 5, 3, 7, 0, 2, 0, 6]], rhs=[[21], [2], [18]],
                              **FULL3)
    assert rc == 0
    np.testing.assert_allclose(x.ravel(), [3, -1, 2], rtol=1e-12)
    rc, _ = ol.block_lu_solve(1, blocks=[[[v]] for v in [0, 1, 5, 3, 7, 0, 2, 0, 6]], rhs=[[21], [2], [18]],
                              **FULL3)
    # first pivot zero -> SparseMatrixError (no pivoting across blocks)
    assert rc == 1


def test_block_sparse_lu_with_fill_ins():
    rc, x = ol.block_lu_solve(2, blocks=BLOCKS_FILL, rhs=[[38, 356], [-389, 2], [44, 611]], **FULL3)
    assert rc == 0
    np.testing.assert_allclose(x, [[3, 4], [-1, -2], [5, 6]], rtol=1e-12, atol=1e-12)
    singular = [b for b in BLOCKS_FILL]
    singular[0] = [[0, 0], [100, 0]]
    rc, _ = ol.block_lu_solve(2, blocks=singular, rhs=[[38, 356], [-389, 2], [44, 611]], **FULL3)
    assert rc == 1


def test_block_sparse_lu_without_fill_ins():
    blocks = [[[0, 200], [3, 1]], [[1, 2], [-3, 4]], [[1, 2], [7, -1]], [[0, 1], [100, 0]], [[3, 4], [5, 6]],
              [[5, 6], [-7, 8]], [[1, 0], [0, 100]]]
    rc, x = ol.block_lu_solve(2, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [0, 3, 6], blocks,
                              [[-389, 2], [38, 356], [44, 611]])
    assert rc == 0
    np.testing.assert_allclose(x, [[-1, -2], [3, 4], [5, 6]], rtol=1e-12, atol=1e-12)


def test_block_sparse_lu_four_node_meshed():
    blocks = [[[20, 1], [2, 21]], [[1, 0], [2, -1]], [[-1, 2], [0, 1]], [[0, 1], [-2, 1]], [[22, -1], [1, 23]],
              [[2, 1], [-1, 0]], [[0, 0], [0, 0]], [[1, -2], [0, 1]], [[24, 2], [-1, 25]], [[-2, 1], [1, 2]],
              [[1, 1], [-1, 2]], [[0, 0], [0, 0]], [[0, -1], [2, 1]], [[26, -2], [1, 27]]]
    rc, x = ol.block_lu_solve(2, [0, 3, 7, 10, 14], [0, 1, 3, 0, 1, 2, 3, 1, 2, 3, 0, 1, 2, 3], [0, 4, 8, 13],
                              blocks, [[21, 40], [-17, 64], [82, -47], [55, 38]])
    assert rc == 0
    np.testing.assert_allclose(x.ravel(), [1, 2, -1, 3, 4, -2, 2, 1], rtol=1e-12, atol=1e-12)


def test_block_lu_6x6_against_dense_solve():
    # mathematics fixes the answer: compare the 6x6 block path with a dense solve of the assembled matrix
    rng = np.random.default_rng(7)
    n = 4
    # loop 0-1-2-3-0: rows 1 and 3 get the fill-in (1,3)/(3,1) created by eliminating 0
    pattern = {0: [0, 1, 3], 1: [0, 1, 2, 3], 2: [1, 2, 3], 3: [0, 1, 2, 3]}
    ip, ci, dg = [0], [], []
    for r in range(n):
        cols = pattern[r]
        ci += cols
        dg.append(ip[-1] + cols.index(r))
        ip.append(len(ci))
    blocks = []
    dense = np.zeros((6 * n, 6 * n))
    k = 0
    for r in range(n):
        for c in ci[ip[r]:ip[r + 1]]:
            b = rng.normal(size=(6, 6)) + (20 * np.eye(6) if r == c else 0)
            blocks.append(b)
            dense[6 * r:6 * r + 6, 6 * c:6 * c + 6] = b
            k += 1
    rhs = rng.normal(size=(n, 6))
    rc, x = ol.block_lu_solve(6, ip, ci, dg, blocks, rhs)
    assert rc == 0
    np.testing.assert_allclose(x.ravel(), np.linalg.solve(dense, rhs.ravel()), rtol=1e-10, atol=1e-12)


# ------------------------------------------------------------------------------------------------
# closed-form 3-bus grid, NR to 1e-12
# ------------------------------------------------------------------------------------------------
def closed_form_grid(p):
    deg = math.pi / 180
    shift = cmath.exp(1j * math.pi / 6)
    u0 = 1.08 * cmath.exp(-1j * deg)
    u1 = 0.97 * cmath.exp(-4j * deg)
    u2 = 0.90 * cmath.exp(-37j * deg)
    y0, ys0, y1 = 1.0 - 2.0j, 0.05 + 0.2j, 3.0 - 4.0j
    b0f = (u0 - u1) * y0 + u0 * ys0
    b0t = (u1 - u0) * y0 + u1 * ys0
    b1f = (u1 - u2 * shift) * y1
    b1t = (u2 - u1 * shift.conjugate()) * y1
    uref, yref = 1.1, 10.0 - 50.0j
    src = yref * (uref - u0)
    i0 = -src + b0f
    i1 = b0t + b1f
    s0 = (i0.conjugate() * u0)
    s1 = (i1.conjugate() * u1)
    ys = -b1t / u2
    s_inj = [s0 / 3, s0 / 3 / 1.08, s0 / 3 / 1.08 ** 2, s1 / 3, s1 / 3 / 0.97, s1 / 3 / 0.97 ** 2, 0.0]
    types = [0, 2, 1, 0, 2, 1, 0]
    if p == 1:
        bp = np.array([[y0 + ys0, -y0, -y0, y0 + ys0], [y1, -y1 * shift, -y1 * shift.conjugate(), y1]])
        bp = bp.reshape(2, 4, 1, 1)
        sp = np.array([ys]).reshape(1, 1, 1)
        si = np.array(s_inj).reshape(7, 1)
    else:
        def tsm(s, m):
            return np.array([[s, m, m], [m, s, m], [m, m, s]], dtype=complex)
        y0_0 = 0.5 + 0.5j
        y0a = tsm(2 * y0 + y0_0, y0_0 - y0) / 3
        ys0a = tsm(ys0, 0)
        y1_1 = tsm(2 * y1, -y1) / 3
        y1_3 = np.array([[-y1, y1, 0], [0, -y1, y1], [y1, 0, -y1]]) / math.sqrt(3)
        bp = np.array([[y0a + ys0a, -y0a, -y0a, y0a + ys0a], [y1_1, y1_3, y1_3.T, y1_1]])
        ys_0 = ys * 0.2
        sp = (tsm(2 * ys + ys_0, ys_0 - ys) / 3).reshape(1, 3, 3)
        si = np.repeat(np.array(s_inj).reshape(7, 1), 3, axis=1)
    grid = dict(n_bus=3, branch_bus=[(0, 1), (1, 2)], branch_param=bp, shunt_bus=[2], shunt_param=sp,
                source_bus=[0], source_y=[[yref, yref]], u_source=[uref], load_bus=[0, 0, 0, 1, 1, 1, 2],
                load_type=types, s_inj=si)
    a2 = cmath.exp(-2j * math.pi / 3)
    rot = np.array([1, a2, a2.conjugate()]) if p == 3 else np.array([1.0])
    ref = dict(u=np.outer([u0, u1, u2], rot), source_i=src * rot, branch_if=np.outer([b0f, b1f], rot),
               branch_it=np.outer([b0t, b1t], rot), shunt_i=b1t * rot,
               load_s=[s0 / 3] * 3 + [s1 / 3] * 3 + [0.0], s0=s0, s1=s1)
    return grid, ref


@pytest.mark.parametrize("p", [1, 3])
def test_closed_form_nr(p):
    grid, ref = closed_form_grid(p)
    it, r = ol.nr_math_model(p, **grid, err_tol=1e-12, max_iter=20)
    # the iteration count is pinned by the dense brute force (same initial guess and stopping rule)
    it_dense, u_dense = dense_nr(p, **grid, err_tol=1e-12, max_iter=20)
    assert it == it_dense == 4
    np.testing.assert_allclose(u_dense, ref["u"], atol=1e-12)
    for tol, expect in ((1e-4, 2), (1e-6, 3), (1e-8, 3), (1e-10, 4)):
        it_t, _ = ol.nr_math_model(p, **grid, err_tol=tol, max_iter=20)
        assert it_t == dense_nr(p, **grid, err_tol=tol, max_iter=20)[0] == expect, tol
    np.testing.assert_allclose(r["u"], ref["u"], atol=1e-12)
    np.testing.assert_allclose(r["source"][0, 1], ref["source_i"], atol=1e-12)
    np.testing.assert_allclose(r["branch"][:, 2], ref["branch_if"], atol=1e-12)
    np.testing.assert_allclose(r["branch"][:, 3], ref["branch_it"], atol=1e-12)
    np.testing.assert_allclose(r["shunt"][0, 1], ref["shunt_i"], atol=1e-12)
    for k in range(7):
        np.testing.assert_allclose(r["load"][k, 0], np.full(p, ref["load_s"][k]), atol=1e-12)


@pytest.mark.parametrize("p", [1, 3])
def test_closed_form_const_z(p):
    grid, ref = closed_form_grid(p)
    s = np.array(grid["s_inj"])
    for i in range(6):
        s[i] = s[i] * 3 if i % 3 == 2 else 0.0
    grid["s_inj"] = s
    it, r = ol.nr_math_model(p, **grid, err_tol=1e-12, max_iter=20)
    assert it >= 1
    np.testing.assert_allclose(r["u"], ref["u"], atol=1e-12)


@pytest.mark.parametrize("p", [1, 3])
def test_closed_form_not_converge_and_singular(p):
    grid, _ = closed_form_grid(p)
    s = np.array(grid["s_inj"])
    s[6] = 1e6
    it, _ = ol.nr_math_model(p, **{**grid, "s_inj": s}, err_tol=1e-12, max_iter=20)
    assert it == -2
    bp = np.zeros_like(np.asarray(grid["branch_param"]))
    sp = np.zeros_like(np.asarray(grid["shunt_param"]))
    it, _ = ol.nr_math_model(p, **{**grid, "branch_param": bp, "shunt_param": sp}, err_tol=1e-12, max_iter=20)
    assert it == -1


@pytest.mark.parametrize("p", [1, 3])
def test_single_iteration_infinite_tolerance(p):
    """test_math_solver_pf.hpp:118-130: err_tol = inf and max_iter = 1 run exactly one iteration without
    IterationDiverge, and the result lies within 0.15 of the exact solution."""
    grid, ref = closed_form_grid(p)
    it, r = ol.nr_math_model(p, **grid, err_tol=math.inf, max_iter=1)
    assert it == 1
    np.testing.assert_allclose(r["u"], ref["u"], atol=0.15)
    it_dense, u_dense = dense_nr(p, **grid, err_tol=math.inf, max_iter=1)
    assert it_dense == 1
    np.testing.assert_allclose(r["u"], u_dense, atol=1e-12)


@pytest.mark.parametrize("p", [1, 3])
def test_max_iter_boundary(p):
    """iterative_pf_solver.hpp:53-58: IterationDiverge fires on entering iteration max_iter + 1. A grid that
    converges in k iterations converges with max_iter = k and diverges with max_iter = k - 1."""
    grid, _ = closed_form_grid(p)
    k, _ = ol.nr_math_model(p, **grid, err_tol=1e-12, max_iter=20)
    assert k == 4
    it, _ = ol.nr_math_model(p, **grid, err_tol=1e-12, max_iter=k)
    assert it == k
    it, _ = ol.nr_math_model(p, **grid, err_tol=1e-12, max_iter=k - 1)
    assert it == -2
    assert dense_nr(p, **grid, err_tol=1e-12, max_iter=k)[0] == k
    assert dense_nr(p, **grid, err_tol=1e-12, max_iter=k - 1)[0] == -2


def _brute_force_grids():
    """100 random grids of 3-50 nodes (sym, asym and asym loads in sym calculations), each under six load
    scales from light to beyond collapse, at three tolerances: iteration counts 1-9 and divergent cases."""
    from power_grid_model_amd import initialize_array

    rng = np.random.default_rng(0)
    for g in range(100):
        sym = g % 4 != 3
        n = int(rng.integers(3, 51))
        inp = random_grid(rng, n, int(rng.integers(0, 4)), sym, n_asym=0 if sym and g % 2 else 2)
        scales = [0.05, 1.0, 4.0, 10.0, 20.0, 40.0]
        tol = [1e-8, 1e-12, 1e-5][g % 3]
        ld = inp["sym_load"]
        upd = initialize_array("update", "sym_load", (len(scales), len(ld)))
        upd["id"] = ld["id"]
        for k, sc in enumerate(scales):
            upd["p_specified"][k] = ld["p_specified"] * sc
            upd["q_specified"][k] = ld["q_specified"] * sc
        yield g, inp, upd, scales, sym, tol


def test_dense_newton_brute_force_iteration_counts():
    """The oracle's per-scenario iteration counts, divergence set and voltages against the dense numpy Newton
    brute force (tests/dense_newton.py) on 600 scenarios of 100 random grids: counts exact, voltages within
    1e-12 pu."""
    seen = set()
    for g, inp, upd, scales, sym, tol in _brute_force_grids():
        out, it, st, _ = ol.power_flow(inp, {"sym_load": upd}, len(scales), symmetric=sym, err_tol=tol)
        for k, sc in enumerate(scales):
            mod = {kk: v.copy() for kk, v in inp.items()}
            mod["sym_load"]["p_specified"] *= sc
            mod["sym_load"]["q_specified"] *= sc
            it_dense, u_dense = dense_nr(1 if sym else 3, **dense_from_input(mod, sym), err_tol=tol)
            it_oracle = int(it[k]) if st[k] == 0 else (-2 if st[k] == 2 else -1)
            assert it_oracle == it_dense, (g, k, it_oracle, it_dense)
            seen.add(it_dense)
            if it_dense > 0:
                u = out["node"]["u_pu"][k] * np.exp(1j * out["node"]["u_angle"][k])
                np.testing.assert_allclose(u.reshape(u_dense.shape), u_dense, atol=1e-12, rtol=0)
    assert {-2, 1, 2, 3, 4, 5, 6}.issubset(seen), seen


def test_two_bus_quadratic_closed_form():
    # source(y_ref) - bus0 - line(y) - bus1 with a const-power load: V1 solves a quadratic.
    # Use an ideal-ish source (huge y_ref) so bus0 = u_ref; then u1 = u0 - z * conj(s / u1).
    z = 0.01 + 0.05j
    s = -(0.5 + 0.2j)  # injection (load)
    grid = dict(n_bus=2, branch_bus=[(0, 1)], branch_param=np.array([[1 / z, -1 / z, -1 / z, 1 / z]]).reshape(
        1, 4, 1, 1), shunt_bus=[], shunt_param=np.zeros((0, 1, 1)), source_bus=[0], source_y=[[1e12, 1e12]],
        u_source=[1.0], load_bus=[1], load_type=[0], s_inj=[[s]])
    it, r = ol.nr_math_model(1, **grid, err_tol=1e-13, max_iter=30)
    u1 = r["u"][1, 0]
    u0 = r["u"][0, 0]
    # residual of the exact power balance at bus 1
    assert abs(u1 * np.conj((u1 - u0) / z) - s) < 1e-9
    # closed form for the real-axis-referenced problem: |u1|^2 - u0* u1 ... check via numpy polynomial root
    assert abs(u0 - 1.0) < 1e-9


def _two_node_grid(load_type):
    from pgm_formats import INPUT, empty

    node = empty(INPUT["node"], 2)
    node["id"] = [1, 2]
    node["u_rated"] = 10.5e3
    line = empty(INPUT["line"], 1)
    line["id"] = 3
    line["from_node"] = 1
    line["to_node"] = 2
    line["from_status"] = 1
    line["to_status"] = 1
    line[["r1", "x1", "c1", "tan1", "r0", "x0", "c0", "tan0", "i_n"]] = (0.5, 1.5, 0.0, 0.0, 1.5, 4.5, 0.0, 0.0, 500.0)
    src = empty(INPUT["source"], 1)
    src["id"] = 4
    src["node"] = 1
    src["status"] = 1
    src["u_ref"] = 1.02
    src["sk"] = 1e9
    load = empty(INPUT["sym_load"], 1)
    load["id"] = 5
    load["node"] = 2
    load["status"] = 1
    load["type"] = load_type
    load["p_specified"] = 2e6
    load["q_specified"] = 0.6e6
    return {"node": node, "line": line, "source": src, "sym_load": load}


@pytest.mark.parametrize("load_type", [0, 1])
def test_linear_method_closed_form(load_type):
    """Linear power flow (linear_pf_solver.hpp): every load is the admittance -conj(s) at 1 pu, one
    complex solve, load results as const_y. A two-node grid solved with numpy."""
    grid = _two_node_grid(load_type)
    out, it, st, msg = ol.power_flow(grid, method="linear")
    assert st[0] == 0
    zb = 10.5e3 ** 2 / 1e6
    y_line = 1.0 / ((0.5 + 1.5j) / zb)
    z = 1e6 / 1e9  # source impedance in pu, rx = 0.1
    x1 = z / math.sqrt(1 + 0.1 ** 2)
    y_src = 1.0 / (0.1 * x1 + 1j * x1)
    s = -(2e6 + 0.6e6j) / 1e6
    y_load = -np.conj(s)
    a = np.array([[y_src + y_line, -y_line], [-y_line, y_line + y_load]])
    u = np.linalg.solve(a, np.array([y_src * 1.02, 0.0]))
    np.testing.assert_allclose(out["node"]["u_pu"][0], np.abs(u), rtol=0, atol=1e-12)
    np.testing.assert_allclose(out["node"]["u_angle"][0], np.angle(u), rtol=0, atol=1e-12)
    # const_y results: p = p_specified |u|^2 whatever the load type
    np.testing.assert_allclose(out["sym_load"]["p"][0, 0], 2e6 * abs(u[1]) ** 2, rtol=1e-12)
    np.testing.assert_allclose(out["sym_load"]["q"][0, 0], 0.6e6 * abs(u[1]) ** 2, rtol=1e-12)
    # Newton-Raphson on an all-const_y grid runs the linear method too (math_solver.hpp:43-64)
    nr = ol.power_flow(grid)[0]
    if load_type == 1:
        assert nr["node"]["u_pu"].tobytes() == out["node"]["u_pu"].tobytes()
    else:
        assert abs(nr["node"]["u_pu"][0, 1] - abs(u[1])) > 1e-6


# ------------------------------------------------------------------------------------------------
# reference validation cases
# ------------------------------------------------------------------------------------------------
CASES = list_cases(tap=False)


@pytest.mark.parametrize("name,sym,batch", CASES, ids=[f"{c[0]}-{'sym' if c[1] else 'asym'}"
                                                       f"{'-batch' if c[2] else ''}" for c in CASES])
def test_validation_case(name, sym, batch):
    params, input_data, update, n_scenarios, expected = load_case(name, sym, batch)
    out, iters, status, msgs = ol.power_flow(input_data, update, n_scenarios, symmetric=sym, threading=-1)
    assert (status == 0).all(), msgs
    for s, exp in enumerate(expected):
        compare_result(out, exp, params["rtol"], params["atol"], s)


METHOD_CASES = [(m, *c) for m in ("linear", "iterative_current", "linear_current") for c in list_cases(m, tap=False)]


@pytest.mark.parametrize("method,name,sym,batch", METHOD_CASES,
                         ids=[f"{c[0]}-{c[1]}-{'sym' if c[2] else 'asym'}{'-batch' if c[3] else ''}"
                              for c in METHOD_CASES])
def test_validation_case_other_methods(method, name, sym, batch):
    params, input_data, update, n_scenarios, expected = load_case(name, sym, batch)
    out, iters, status, msgs = ol.power_flow(input_data, update, n_scenarios, symmetric=sym, method=method)
    assert (status == 0).all(), msgs
    for s, exp in enumerate(expected):
        compare_result(out, exp, params["rtol"], params["atol"], s)


TAP_CASES = [(m, *c) for m in ("newton_raphson", "iterative_current") for c in list_cases(m, tap=True)]


def _with_regulator_output(out, taps, input_data):
    reg = input_data["transformer_tap_regulator"]
    rec = np.zeros(taps.shape, dtype=SYM_OUTPUT["transformer_tap_regulator"])
    rec["id"] = reg["id"][None, :]
    rec["tap_pos"] = taps
    rec["energized"] = (taps != -128).astype(np.int8)
    return dict(out, transformer_tap_regulator=rec)


@pytest.mark.parametrize("method,name,sym,batch", TAP_CASES,
                         ids=[f"{c[0]}-{c[1]}-{'sym' if c[2] else 'asym'}{'-batch' if c[3] else ''}"
                              for c in TAP_CASES])
def test_validation_case_tap_changer(method, name, sym, batch):
    """The oracle's automatic tap changer against the reference's tap regulator cases: every output and the
    regulated tap positions."""
    params, input_data, update, n_scenarios, expected = load_case(name, sym, batch)
    out, taps, iters, status, msgs = ol.power_flow_tap(input_data, update, n_scenarios, symmetric=sym,
                                                       strategy=params["tap_changing_strategy"], method=method)
    assert (status == 0).all(), msgs
    out = _with_regulator_output(out, taps, input_data)
    for s, exp in enumerate(expected):
        compare_result(out, exp, params["rtol"], params["atol"], s)


TAP_RAISE_CASES = list_cases(tap=True, raises=True)


@pytest.mark.parametrize("name,sym,batch", TAP_RAISE_CASES, ids=[c[0] for c in TAP_RAISE_CASES])
def test_validation_case_tap_changer_raises(name, sym, batch):
    """Cases the reference reports as MaxIterationReached."""
    params, input_data, update, n_scenarios, _ = load_case(name, sym, batch)
    _, _, _, status, msgs = ol.power_flow_tap(input_data, update, n_scenarios, symmetric=sym,
                                              strategy=params["tap_changing_strategy"])
    assert (status == 4).all() and all(m.startswith("Maximum number of iterations reached!") for m in msgs), msgs


def test_iterative_current_reaches_newton_raphson_fixed_point():
    """Iterative current and Newton-Raphson solve the same equations: at a tight tolerance they agree
    to it, and linear current (one iteration) does not."""
    grid = _two_node_grid(0)
    nr = ol.power_flow(grid, err_tol=1e-12)[0]
    ic, it, st, _ = ol.power_flow(grid, err_tol=1e-12, max_iter=100, method="iterative_current")
    assert st[0] == 0 and it[0] > 2
    np.testing.assert_allclose(ic["node"]["u_pu"], nr["node"]["u_pu"], atol=1e-11)
    np.testing.assert_allclose(ic["sym_load"]["p"], nr["sym_load"]["p"], rtol=1e-10)
    lc, it1, _, _ = ol.power_flow(grid, method="linear_current")
    assert it1[0] == 1 and abs(lc["node"]["u_pu"][0, 1] - nr["node"]["u_pu"][0, 1]) > 1e-9


def test_batch_threading_identical():
    name = "dummy-test-batch-newton"
    params, input_data, update, n, expected = load_case(name, True, True)
    ref = ol.power_flow(input_data, update, n, threading=-1)[0]
    for thr in (0, 1, 2):
        out = ol.power_flow(input_data, update, n, threading=thr)[0]
        for comp in ref:
            assert out[comp].tobytes() == ref[comp].tobytes()
