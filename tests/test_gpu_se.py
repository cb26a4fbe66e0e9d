"""GPU state estimation: the HIP Newton-Raphson state estimation against the reference's state-estimation
validation cases (tests/golden/state_estimation/**, copied unchanged from the reference's tests/data), compared
with the reference's own rule and tolerances (params.json rtol / atol, residual patterns included)."""

import numpy as np
import pytest
from validation_utils import GOLDEN_SE, compare_result, list_cases, load_case

pgm = pytest.importorskip("power_grid_model_amd")
from power_grid_model_amd import PowerGridModel, synthetic  # noqa: E402

pytestmark = pytest.mark.gpu

SE_CASES = [c for c in list_cases("newton_raphson", root=GOLDEN_SE) if c[1]]


@pytest.mark.parametrize("name,sym,batch", SE_CASES,
                         ids=[f"{c[0]}{'-batch' if c[2] else ''}" for c in SE_CASES])
def test_state_estimation_case_gpu(name, sym, batch):
    params, input_data, update, n_scenarios, expected = load_case(name, sym, batch, root=GOLDEN_SE)
    model = PowerGridModel(input_data)
    out = model.calculate_state_estimation(symmetric=True, update_data=update, calculation_method="newton_raphson")
    if not batch:
        out = {k: v.reshape(1, *v.shape) for k, v in out.items()}
    for s, exp in enumerate(expected):
        compare_result(out, exp, params["rtol"], params["atol"], s)


def test_state_estimation_refuses_other_methods():
    """The reference's default state-estimation method (iterative_linear) is not part of this build: the
    request fails loudly instead of running another method."""
    params, input_data, update, n_scenarios, expected = load_case("1os2msr", True, False, root=GOLDEN_SE)
    model = PowerGridModel(input_data)
    with pytest.raises(pgm.PowerGridError, match="newton_raphson"):
        model.calculate_state_estimation(calculation_method="iterative_linear")


def test_state_estimation_batch_lanes_independent():
    """A batch of the same grid with perturbed measurements: every scenario equals the single calculation of
    its own measurements (the lanes of a group share only the structure)."""
    params, input_data, update, n_scenarios, expected = load_case("1os2msr", True, False, root=GOLDEN_SE)
    rng = np.random.default_rng(3)
    n = 70
    vs = input_data["sym_voltage_sensor"]
    upd = np.zeros((n, len(vs)), dtype=pgm.power_grid_meta_data["update"]["sym_voltage_sensor"].dtype)
    upd["id"] = vs["id"]
    upd["u_sigma"] = np.nan
    upd["u_angle_measured"] = np.nan
    upd["u_measured"] = vs["u_measured"] * (1.0 + 1e-3 * rng.standard_normal((n, len(vs))))
    model = PowerGridModel(input_data)
    batch = model.calculate_state_estimation(update_data={"sym_voltage_sensor": upd},
                                             calculation_method="newton_raphson")
    for s in (0, 33, 69):
        data = {k: v.copy() for k, v in input_data.items()}
        data["sym_voltage_sensor"]["u_measured"] = upd["u_measured"][s]
        single = PowerGridModel(data).calculate_state_estimation(calculation_method="newton_raphson")
        for comp in ("node", "line"):
            for attr in single[comp].dtype.names:
                np.testing.assert_allclose(batch[comp][attr][s], single[comp][attr], rtol=1e-12, atol=1e-12,
                                           err_msg=f"{comp}.{attr} scenario {s}")


SE_ORACLE_CASES = [
    ("radial_feeder_60", lambda: synthetic.radial_feeder_case(n_node=60, n_scenarios=70, seed=5)),
    ("radial_200", lambda: synthetic.radial_200(n_scenarios=6, seed=6)),
]


@pytest.mark.parametrize("name,make", SE_ORACLE_CASES, ids=[c[0] for c in SE_ORACLE_CASES])
def test_state_estimation_against_oracle(name, make):
    """Synthetic grids with noisy measurements taken from their power flows (voltage magnitudes on a share of
    the nodes, the source's voltage phasor, every load and source power, from-side flows on a share of the
    lines), a batch of scenarios: the GPU estimate equals the oracle's restatement of the reference's
    Newton-Raphson state estimation scenario by scenario (FP64 both; the same iteration, so agreement far below
    the 1e-8 convergence tolerance)."""
    import oracle_lib

    case = make()
    pf, _, st, msg = oracle_lib.power_flow(case.input, case.update, case.n_scenarios, err_tol=1e-12)
    assert (st == 0).all(), msg
    se_case = synthetic.state_estimation_case(case, pf, voltage_share=0.1, flow_share=0.2)
    want, iters, st, msg = oracle_lib.state_estimation(se_case.input, se_case.update, se_case.n_scenarios)
    assert (st == 0).all(), msg
    got = PowerGridModel(se_case.input).calculate_state_estimation(update_data=se_case.update,
                                                                   calculation_method="newton_raphson")
    upd = se_case.update
    measured = {"sym_voltage_sensor": upd["sym_voltage_sensor"]["u_measured"],
                "sym_power_sensor": np.concatenate([upd["sym_power_sensor"]["p_measured"],
                                                    upd["sym_power_sensor"]["q_measured"]])}
    np.testing.assert_allclose(got["node"]["u_pu"], want["node"]["u_pu"], rtol=1e-9, atol=0)
    np.testing.assert_allclose(got["node"]["u_angle"], want["node"]["u_angle"], rtol=0, atol=1e-9)
    for comp, attrs in (("node", ("p", "q")), ("line", ("p_from", "q_from", "p_to", "q_to")),
                        ("sym_load", ("p", "q")), ("source", ("p", "q")),
                        ("sym_voltage_sensor", ("u_residual",)), ("sym_power_sensor", ("p_residual", "q_residual"))):
        for attr in attrs:
            # a residual is a difference of two values of the measured quantity's size: its absolute bound
            # follows that size
            ref = measured[comp] if comp in measured else want[comp][attr]
            scale = max(float(np.abs(ref).max()), 1.0)
            np.testing.assert_allclose(got[comp][attr], want[comp][attr], rtol=1e-7, atol=1e-9 * scale,
                                       err_msg=f"{comp}.{attr}")
