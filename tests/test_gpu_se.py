"""GPU state estimation: the HIP Newton-Raphson state estimation against the reference's state-estimation
validation cases (tests/golden/state_estimation/**, copied unchanged from the reference's tests/data), compared
with the reference's own rule and tolerances (params.json rtol / atol, residual patterns included)."""

import numpy as np
import pytest
from validation_utils import GOLDEN_SE, compare_result, list_cases, load_case

pgm = pytest.importorskip("power_grid_model_amd")
from power_grid_model_amd import PowerGridModel  # noqa: E402

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
