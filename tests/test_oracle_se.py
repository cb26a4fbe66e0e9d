"""The oracle's Newton-Raphson state estimation against the reference's state-estimation validation cases
(tests/golden/state_estimation/**, copied unchanged from the reference's tests/data), with the reference's own
comparison rule and tolerances (params.json rtol / atol). This pins the CPU restatement that the GPU state
estimation is checked against beyond those cases."""

import numpy as np
import oracle_lib
import pytest
from validation_utils import GOLDEN_SE, compare_result, list_cases, load_case

SE_CASES = [c for c in list_cases("newton_raphson", root=GOLDEN_SE) if c[1]]


@pytest.mark.parametrize("name,sym,batch", SE_CASES,
                         ids=[f"{c[0]}{'-batch' if c[2] else ''}" for c in SE_CASES])
def test_oracle_state_estimation_case(name, sym, batch):
    params, input_data, update, n_scenarios, expected = load_case(name, sym, batch, root=GOLDEN_SE)
    out, iters, status, messages = oracle_lib.state_estimation(input_data, update, n_scenarios)
    assert list(status) == [0] * n_scenarios, messages
    for s, exp in enumerate(expected):
        compare_result(out, exp, params["rtol"], params["atol"], s)


def test_oracle_state_estimation_planted_mistake_fails():
    """A measured value moved by 1 % changes the estimate beyond the case's tolerance: the comparison above is
    sensitive to the measurement model, not only to the solver converging."""
    params, input_data, update, n_scenarios, expected = load_case("1os2msr", True, False, root=GOLDEN_SE)
    data = {k: v.copy() for k, v in input_data.items()}
    data["sym_voltage_sensor"]["u_measured"] *= 1.01
    out, iters, status, messages = oracle_lib.state_estimation(data)
    assert status[0] == 0, messages
    with pytest.raises(AssertionError):
        compare_result(out, expected[0], params["rtol"], params["atol"], 0)


def test_oracle_state_estimation_exact_measurements_reproduce_power_flow():
    """Measurements taken from a converged power flow (every node voltage and every appliance power, exact)
    give back that power flow's state: the estimator's fixed point is the true state when the measurements are
    consistent, independently of their variances."""
    from pgm_formats import INPUT

    nodes = np.zeros(4, dtype=INPUT["node"])
    nodes["id"] = [1, 2, 3, 4]
    nodes["u_rated"] = 10e3
    lines = np.zeros(3, dtype=INPUT["line"])
    lines["id"] = [5, 6, 7]
    lines["from_node"] = [1, 2, 2]
    lines["to_node"] = [2, 3, 4]
    lines["from_status"] = lines["to_status"] = 1
    lines["r1"], lines["x1"], lines["c1"], lines["tan1"] = 0.2, 0.4, 1e-7, 0.0
    lines["r0"], lines["x0"], lines["c0"], lines["tan0"] = 0.6, 1.2, 1e-7, 0.0
    lines["i_n"] = 500.0
    src = np.zeros(1, dtype=INPUT["source"])
    src["id"], src["node"], src["status"], src["u_ref"], src["u_ref_angle"] = 8, 1, 1, 1.02, 0.0
    src["sk"], src["rx_ratio"], src["z01_ratio"] = 1e20, 0.1, 1.0
    loads = np.zeros(3, dtype=INPUT["sym_load"])
    loads["id"] = [9, 10, 11]
    loads["node"] = [2, 3, 4]
    loads["status"] = 1
    loads["type"] = 0
    loads["p_specified"] = [1e6, 0.5e6, 0.8e6]
    loads["q_specified"] = [0.2e6, 0.1e6, 0.3e6]
    grid = {"node": nodes, "line": lines, "source": src, "sym_load": loads}
    pf, _, st, msg = oracle_lib.power_flow(grid, err_tol=1e-12)
    assert st[0] == 0, msg
    vs = np.zeros(4, dtype=INPUT["sym_voltage_sensor"])
    vs["id"] = [20, 21, 22, 23]
    vs["measured_object"] = [1, 2, 3, 4]
    vs["u_sigma"] = [1.0, 2.0, 3.0, 4.0]
    vs["u_measured"] = pf["node"]["u"][0]
    vs["u_angle_measured"] = pf["node"]["u_angle"][0]
    ps = np.zeros(3, dtype=INPUT["sym_power_sensor"])
    ps["id"] = [30, 31, 32]
    ps["measured_object"] = [9, 10, 11]
    ps["measured_terminal_type"] = 4  # load
    ps["power_sigma"] = [1e3, 2e3, 5e3]
    ps["p_measured"] = pf["sym_load"]["p"][0]
    ps["q_measured"] = pf["sym_load"]["q"][0]
    ps["p_sigma"] = ps["q_sigma"] = np.nan
    se, iters, st, msg = oracle_lib.state_estimation(dict(grid, sym_voltage_sensor=vs, sym_power_sensor=ps),
                                                     err_tol=1e-12)
    assert st[0] == 0, msg
    np.testing.assert_allclose(se["node"]["u_pu"], pf["node"]["u_pu"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(se["node"]["u_angle"], pf["node"]["u_angle"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(se["line"]["p_from"], pf["line"]["p_from"], rtol=1e-7, atol=1e-2)
    np.testing.assert_allclose(se["sym_voltage_sensor"]["u_residual"], 0.0, atol=1e-9)
    np.testing.assert_allclose(se["sym_power_sensor"]["p_residual"], 0.0, atol=1e-2)


def test_oracle_state_estimation_sparse_equals_dense():
    """The oracle's two solutions of the gain system agree on every reference case: the block sparse LU on the
    Y-bus pattern with the reference's pivot perturbation and iterative refinement, and a dense LU with partial
    pivoting of the same matrix. Some cases perturb pivots (the reference marks them possibly ill-conditioned),
    so the perturbation and the refinement are part of what the cases pin."""
    oracle_lib.se_counters(reset=True)
    worst = 0.0
    try:
        for name, sym, batch in SE_CASES:
            params, input_data, update, n_scenarios, expected = load_case(name, sym, batch, root=GOLDEN_SE)
            oracle_lib.set_se_dense(False)
            sparse, _, st_s, msg = oracle_lib.state_estimation(input_data, update, n_scenarios)
            oracle_lib.set_se_dense(True)
            dense, _, st_d, _ = oracle_lib.state_estimation(input_data, update, n_scenarios)
            assert list(st_s) == list(st_d) == [0] * n_scenarios, (name, msg)
            d = np.abs(sparse["node"]["u_pu"] - dense["node"]["u_pu"]).max()
            worst = max(worst, d)
            assert d < 1e-9, name
    finally:
        oracle_lib.set_se_dense(False)
    perturbed, solves = oracle_lib.se_counters(reset=True)
    assert perturbed > 0 and solves >= 2 * perturbed, (perturbed, solves)
