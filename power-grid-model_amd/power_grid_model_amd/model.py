"""PowerGridModel with the reference's power-flow interface (src/power_grid_model/_core/power_grid_model.py:
PowerGridModel(input_data, system_frequency), calculate_power_flow(...), update, copy, get_indexer),
running every calculation on the MI355X through libpower_grid_model_c.so."""

from __future__ import annotations

import ctypes as C
from enum import IntEnum

import numpy as np

from ._lib import Handle, lib
from .errors import PowerGridBatchError, error_class, raise_regular_error
from .meta import initialize_array, power_grid_meta_data


class CalculationType(IntEnum):
    power_flow = 0
    state_estimation = 1
    short_circuit = 2


class CalculationMethod(IntEnum):
    default_method = -128
    linear = 0
    newton_raphson = 1
    iterative_linear = 2
    iterative_current = 3
    linear_current = 4
    iec60909 = 5


class ComponentAttributeFilterOptions(IntEnum):
    """Columnar output of every attribute of the selected components (enum.py:128-134)."""

    everything = 0
    relevant = 1


class LoadGenType(IntEnum):
    const_power = 0
    const_impedance = 1
    const_current = 2


class WindingType(IntEnum):
    wye = 0
    wye_n = 1
    delta = 2
    zigzag = 3
    zigzag_n = 4


class BranchSide(IntEnum):
    from_side = 0
    to_side = 1


class FactorPrecision(IntEnum):
    """Precision of the Newton-Raphson factorisation (extension of this build)."""

    fp64 = 0
    fp32 = 1


def _enum_value(v, enum_type):
    if isinstance(v, str):
        return enum_type[v].value
    return int(v)


class _Dataset:
    """Keeps a PGM dataset handle and the numpy buffers it points to alive."""

    def __init__(self, ptr, destroy, keep):
        self.ptr = ptr
        self._destroy = destroy
        self._keep = keep

    def __del__(self):
        if getattr(self, "ptr", None):
            self._destroy(self.ptr)
            self.ptr = None


def _component_buffer(value, dtype, is_batch):
    """Normalise one component of a dataset to (data arrays, elements_per_scenario, total, indptr, batch)."""
    if isinstance(value, dict) and "indptr" in value:
        indptr = np.ascontiguousarray(value["indptr"], dtype=np.int64)
        data = value["data"]
        batch = len(indptr) - 1
        if isinstance(data, np.ndarray):
            data = np.ascontiguousarray(data, dtype=dtype)
            return data, -1, len(data), indptr, batch
        cols = {k: np.ascontiguousarray(v) for k, v in data.items()}
        total = len(next(iter(cols.values()))) if cols else 0
        return cols, -1, total, indptr, batch
    if isinstance(value, np.ndarray):
        arr = np.ascontiguousarray(value, dtype=dtype)
        if is_batch:
            if arr.ndim == 1:
                arr = arr.reshape(1, -1)
            return arr, arr.shape[1], arr.size, None, arr.shape[0]
        return arr, arr.shape[0], arr.shape[0], None, 1
    # columnar: dict of attribute arrays
    cols = {k: np.ascontiguousarray(v) for k, v in value.items()}
    first = next(iter(cols.values()))
    if is_batch:
        if first.ndim == 1:
            cols = {k: v.reshape(1, *v.shape) for k, v in cols.items()}
            first = next(iter(cols.values()))
        return cols, first.shape[1], first.shape[0] * first.shape[1], None, first.shape[0]
    return cols, first.shape[0], first.shape[0], None, 1


def _make_const_dataset(h: Handle, data: dict, dataset_type: str, is_batch: bool) -> _Dataset:
    meta = power_grid_meta_data[dataset_type]
    keep = []
    batch_size = 1
    prepared = []
    for comp, value in data.items():
        if comp not in meta:
            raise_regular_error(f"component {comp} is not supported by this build")
        arrs, eps, total, indptr, batch = _component_buffer(value, meta[comp].dtype, is_batch)
        batch_size = batch if is_batch else 1
        prepared.append((comp, arrs, eps, total, indptr))
    if is_batch and not prepared:
        batch_size = 1
    ptr = lib.PGM_create_dataset_const(h.ptr, dataset_type.encode(), int(is_batch), batch_size)
    h.check()
    ds = _Dataset(ptr, lib.PGM_destroy_dataset_const, keep)
    for comp, arrs, eps, total, indptr in prepared:
        keep.append(arrs)
        keep.append(indptr)
        if isinstance(arrs, dict):
            lib.PGM_dataset_const_add_buffer(h.ptr, ptr, comp.encode(), eps, total,
                                             None if indptr is None else indptr.ctypes.data, None)
            h.check()
            for attr, col in arrs.items():
                lib.PGM_dataset_const_add_attribute_buffer(h.ptr, ptr, comp.encode(), attr.encode(), col.ctypes.data)
                h.check()
        else:
            lib.PGM_dataset_const_add_buffer(h.ptr, ptr, comp.encode(), eps, total,
                                             None if indptr is None else indptr.ctypes.data, arrs.ctypes.data)
            h.check()
    ds.batch_size = batch_size
    return ds


def _make_update(h: Handle, update_data) -> tuple[_Dataset, int]:
    """The batch update dataset of a calculation. A list of update datasets is a cartesian product of batches
    (power_grid_model.py calculate_power_flow update_data list; model.cpp:285-334 of the reference C API): the
    datasets are chained as dimensions, the first dimension outermost, and the batch size is the product of
    their sizes. Returns the first dataset (which keeps the others alive) and the total batch size."""
    if not isinstance(update_data, (list, tuple)):
        ds = _make_const_dataset(h, update_data, "update", True)
        return ds, ds.batch_size
    if not update_data:
        raise_regular_error("update_data is an empty list")
    dims = [_make_const_dataset(h, u, "update", True) for u in update_data]
    for a, b in zip(dims[:-1], dims[1:]):
        lib.PGM_dataset_const_set_next_cartesian_product_dimension(h.ptr, a.ptr, b.ptr)
        h.check()
    dims[0]._keep.append(dims[1:])
    total = 1
    for d in dims:
        total *= d.batch_size
    return dims[0], total


def _make_mutable_dataset(h: Handle, data: dict, dataset_type: str, is_batch: bool, batch_size: int) -> _Dataset:
    ptr = lib.PGM_create_dataset_mutable(h.ptr, dataset_type.encode(), int(is_batch), batch_size)
    h.check()
    keep = []
    ds = _Dataset(ptr, lib.PGM_destroy_dataset_mutable, keep)
    for comp, arr in data.items():
        keep.append(arr)
        if isinstance(arr, dict):
            # columnar: one buffer per requested attribute (dataset.h add_attribute_buffer); the element count is
            # the axis after the batch axis (asymmetric attributes carry a trailing phase axis)
            first = next(iter(arr.values())) if arr else None
            n = 0 if first is None else (first.shape[1] if is_batch else first.shape[0])
            lib.PGM_dataset_mutable_add_buffer(h.ptr, ptr, comp.encode(), n, n * batch_size, None, None)
            h.check()
            for attr, col in arr.items():
                lib.PGM_dataset_mutable_add_attribute_buffer(h.ptr, ptr, comp.encode(), attr.encode(),
                                                             col.ctypes.data)
                h.check()
            continue
        n = arr.shape[1] if is_batch else arr.shape[0]
        lib.PGM_dataset_mutable_add_buffer(h.ptr, ptr, comp.encode(), n, n * batch_size, None, arr.ctypes.data)
        h.check()
    return ds


def _process_output_filter(out_type: str, data_filter, available: list[str]) -> dict:
    """Normalise output_component_types as the reference does (utils.py:456-486): None = every
    component row-based, a list or set = those components row-based, a filter option = every component
    columnar, a dict = per component None (row-based), an attribute list or a filter option (columnar)."""
    if data_filter is None:
        flt = {c: None for c in available}
    elif isinstance(data_filter, ComponentAttributeFilterOptions):
        flt = {c: data_filter for c in available}
    elif isinstance(data_filter, (list, set)):
        flt = {c: None for c in data_filter}
    elif isinstance(data_filter, dict) and all(
            a is None or isinstance(a, (set, list, ComponentAttributeFilterOptions)) for a in data_filter.values()):
        flt = dict(data_filter)
    else:
        raise ValueError(f"Invalid filter provided: {data_filter}")
    meta = power_grid_meta_data[out_type]
    unknown = [c for c in flt if c not in meta]
    if unknown:
        raise KeyError(f"You have specified some unknown component types: {unknown}")
    for c, attrs in flt.items():
        if isinstance(attrs, (list, set)):
            bad = [a for a in attrs if a not in meta[c].dtype.names]
            if bad:
                raise KeyError(f"You have specified some unknown attributes: {{'{c}': {bad}}}")
    return flt


class PowerGridModel:
    """Power grid model on the MI355X (power flow, Newton-Raphson)."""

    def __init__(self, input_data: dict, system_frequency: float = 50.0):
        self._handle = Handle()
        self._ptr = None
        if isinstance(input_data, PowerGridModel):
            self._ptr = lib.PGM_copy_model(self._handle.ptr, input_data._ptr)
            self._handle.check()
            self._all_count = dict(input_data._all_count)
            return
        ds = _make_const_dataset(self._handle, input_data, "input", False)
        self._ptr = lib.PGM_create_model(self._handle.ptr, system_frequency, ds.ptr)
        self._handle.check()
        self._all_count = {k: len(v) if isinstance(v, np.ndarray) else len(next(iter(v.values())))
                           for k, v in input_data.items()}
        self.batch_error: PowerGridBatchError | None = None

    def __del__(self):
        if getattr(self, "_ptr", None):
            lib.PGM_destroy_model(self._ptr)
            self._ptr = None

    def copy(self) -> "PowerGridModel":
        return PowerGridModel(self)

    def make_resident(self, update_data: dict) -> int:
        """Copy the load_gen update buffers of a batch update dataset to HBM once: later calculations on
        this model over the same arrays (same memory, unchanged) read the device copies instead of uploading
        them again. The model keeps a reference to the arrays so that their memory is not reused while the
        device copies exist. Returns the bytes uploaded. Extension of this build (no reference counterpart)."""
        ds = _make_const_dataset(self._handle, update_data, "update", True)
        n = lib.PGM_amd_model_make_resident(self._handle.ptr, self._ptr, ds.ptr)
        self._handle.check()
        self.__dict__.setdefault("_resident_keep", []).append(ds._keep)
        return int(n)

    def autotune(self, update_data: dict, *, symmetric: bool = True, error_tolerance: float = 1e-8,
                 max_iterations: int = 20,
                 calculation_method: CalculationMethod | str = CalculationMethod.newton_raphson,
                 factor_precision: FactorPrecision | str = FactorPrecision.fp64) -> dict:
        """Pick the schedule parameters of this batch shape by measurement: the batch is built and run with
        candidate parameters, and the fastest are kept for later calculations on this model with the same
        shape (number of scenarios, symmetry, factor precision, method). Extension of this build.
        Returns {"params": {...}, "trials": [(params, device ms), ...]}."""
        h = self._handle
        opt = self._options(symmetric, error_tolerance, max_iterations, calculation_method, -1, factor_precision)
        upd = _make_const_dataset(h, update_data, "update", True)
        params = np.zeros(3, dtype=np.int64)
        trials = np.zeros((64, 4), dtype=np.float64)
        n = lib.PGM_amd_autotune(h.ptr, self._ptr, opt.ptr, upd.ptr, params.ctypes.data, trials.ctypes.data, 64)
        h.check()
        names = ("plain_waves", "union_tasks", "packed_target")
        return {"params": dict(zip(names, params.tolist())),
                "trials": [(dict(zip(names, (int(x) for x in t[:3]))), float(t[3])) for t in trials[:min(n, 64)]]}

    def release_device_memory(self) -> None:
        """Free the resident update buffers and the device memory kept from earlier calculations."""
        lib.PGM_amd_model_release_device_memory(self._ptr)
        self.__dict__.pop("_resident_keep", None)

    @property
    def all_component_count(self) -> dict[str, int]:
        return {k: v for k, v in self._all_count.items() if v > 0}

    def update(self, *, update_data: dict):
        ds = _make_const_dataset(self._handle, update_data, "update", False)
        lib.PGM_update_model(self._handle.ptr, self._ptr, ds.ptr)
        self._handle.check()

    def get_indexer(self, component_type: str, ids: np.ndarray) -> np.ndarray:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.empty(ids.shape, dtype=np.int64)
        lib.PGM_get_indexer(self._handle.ptr, self._ptr, component_type.encode(), ids.size, ids.ctypes.data,
                            out.ctypes.data)
        self._handle.check()
        return out

    def _options(self, symmetric, error_tolerance, max_iterations, calculation_method, threading,
                 factor_precision, tap_changing_strategy=0, adjoint_layout=False, refine_steps=False,
                 calculation_type=None):
        h = self._handle
        opt = lib.PGM_create_options(h.ptr)
        lib.PGM_set_calculation_type(h.ptr, opt, CalculationType.power_flow if calculation_type is None
                                     else calculation_type)
        lib.PGM_set_calculation_method(h.ptr, opt, _enum_value(calculation_method, CalculationMethod))
        lib.PGM_set_symmetric(h.ptr, opt, int(bool(symmetric)))
        lib.PGM_set_err_tol(h.ptr, opt, float(error_tolerance))
        lib.PGM_set_max_iter(h.ptr, opt, int(max_iterations))
        lib.PGM_set_threading(h.ptr, opt, int(threading))
        lib.PGM_amd_set_factor_precision(h.ptr, opt, _enum_value(factor_precision, FactorPrecision))
        lib.PGM_set_tap_changing_strategy(h.ptr, opt, _enum_value(tap_changing_strategy, TapChangingStrategy))
        lib.PGM_amd_set_adjoint_layout(h.ptr, opt, int(bool(adjoint_layout)))
        lib.PGM_amd_set_step_refinement(h.ptr, opt, int(bool(refine_steps)))
        h.check()
        return _Dataset(opt, lib.PGM_destroy_options, [])

    def _output_arrays(self, output_component_types, symmetric, is_batch, batch_size):
        """Output buffers as the reference creates them (data_handling.py:146-200): structured arrays for
        row-based components, {attribute: array} for columnar ones."""
        out_type = "sym_output" if symmetric else "asym_output"
        comps = self.all_component_count
        flt = _process_output_filter(out_type, output_component_types, list(comps))
        out = {}
        for c, n in comps.items():
            if c not in flt:
                continue
            shape = (batch_size, n) if is_batch else n
            req = flt[c]
            dtype = power_grid_meta_data[out_type][c].dtype
            if req is None:
                out[c] = initialize_array(out_type, c, shape, empty=True)
            elif isinstance(req, ComponentAttributeFilterOptions):
                out[c] = {a: np.empty(shape, dtype=dtype[a]) for a in dtype.names}
            else:
                out[c] = {a: np.empty(shape, dtype=dtype[a]) for a in req}
        return out_type, out

    def calculate_power_flow(
        self,
        *,
        symmetric: bool = True,
        error_tolerance: float = 1e-8,
        max_iterations: int = 20,
        calculation_method: CalculationMethod | str = CalculationMethod.newton_raphson,
        update_data: dict | list[dict] | None = None,
        threading: int = -1,
        output_component_types=None,
        continue_on_batch_error: bool = False,
        decode_error: bool = True,
        tap_changing_strategy=0,
        factor_precision: FactorPrecision | str = FactorPrecision.fp64,
    ) -> dict:
        """Newton-Raphson power flow of the model, or of every scenario of update_data (a batch).

        tap_changing_strategy other than disabled adjusts the tap position of every transformer with an
        active transformer_tap_regulator until the regulated voltages lie in their bands
        (power_grid_model.py:564-670, optimizer/tap_position_optimizer.hpp)."""
        h = self._handle
        self.batch_error = None
        opt = self._options(symmetric, error_tolerance, max_iterations, calculation_method, threading,
                            factor_precision, tap_changing_strategy)
        is_batch = update_data is not None
        upd = None
        batch_size = 1
        if is_batch:
            upd, batch_size = _make_update(h, update_data)
        out_type, output = self._output_arrays(output_component_types, symmetric, is_batch, batch_size)
        out_ds = _make_mutable_dataset(h, output, out_type, is_batch, batch_size)
        lib.PGM_calculate(h.ptr, self._ptr, opt.ptr, out_ds.ptr, None if upd is None else upd.ptr)
        self._handle_errors(continue_on_batch_error, batch_size, decode_error)
        return output

    def calculate_state_estimation(
        self,
        *,
        symmetric: bool = True,
        error_tolerance: float = 1e-8,
        max_iterations: int = 20,
        calculation_method: CalculationMethod | str = CalculationMethod.newton_raphson,
        update_data: dict | None = None,
        threading: int = -1,
        output_component_types=None,
        continue_on_batch_error: bool = False,
        decode_error: bool = True,
    ) -> dict:
        """Newton-Raphson state estimation of the model, or of every scenario of update_data (a batch), from
        its sym_voltage_sensor and sym_power_sensor measurements (power_grid_model.py calculate_state_estimation,
        math_solver/newton_raphson_se_solver.hpp). Symmetric, method newton_raphson; the reference's default
        method (iterative_linear) is not part of this build and is refused."""
        h = self._handle
        self.batch_error = None
        opt = self._options(symmetric, error_tolerance, max_iterations, calculation_method, threading,
                            FactorPrecision.fp64, calculation_type=CalculationType.state_estimation)
        is_batch = update_data is not None
        upd = None
        batch_size = 1
        if is_batch:
            upd, batch_size = _make_update(h, update_data)
        out_type, output = self._output_arrays(output_component_types, symmetric, is_batch, batch_size)
        out_ds = _make_mutable_dataset(h, output, out_type, is_batch, batch_size)
        lib.PGM_calculate(h.ptr, self._ptr, opt.ptr, out_ds.ptr, None if upd is None else upd.ptr)
        self._handle_errors(continue_on_batch_error, batch_size, decode_error)
        return output

    def _handle_errors(self, continue_on_batch_error: bool, batch_size: int, decode_error: bool):
        h = self._handle
        code = h.code()
        if code == 0:
            return
        msg = h.message()
        if code != 2:
            raise_regular_error(msg)
        n_failed = lib.PGM_n_failed_scenarios(h.ptr)
        failed_ptr = lib.PGM_failed_scenarios(h.ptr)
        errs_ptr = lib.PGM_batch_errors(h.ptr)
        failed = np.array([failed_ptr[i] for i in range(n_failed)], dtype=np.int64)
        messages = [errs_ptr[i].decode() for i in range(n_failed)]
        err = PowerGridBatchError(msg)
        err.failed_scenarios = failed
        err.succeeded_scenarios = np.setdiff1d(np.arange(batch_size), failed)
        err.error_messages = messages
        err.errors = [error_class(m)(m) for m in messages] if decode_error else []
        if continue_on_batch_error:
            self.batch_error = err
            return
        raise err


class TapChangingStrategy(IntEnum):
    disabled = 0
    any_valid_tap = 1
    min_voltage_tap = 2
    max_voltage_tap = 3
    fast_any_tap = 4


class DeviceBatch:
    """A batch kept in HBM: created once (topology grouping, plans, upload), run many times.

    Extension of this build for callers that re-run the same scenarios (bench, sensitivity sweeps)."""

    def __init__(self, model: PowerGridModel, update_data: dict | None, *, symmetric: bool = True,
                 error_tolerance: float = 1e-8, max_iterations: int = 20,
                 factor_precision: FactorPrecision | str = FactorPrecision.fp64,
                 calculation_method: CalculationMethod | str = CalculationMethod.newton_raphson,
                 adjoint_layout: bool = False, refine_steps: bool = False):
        """adjoint_layout: keep every topology group a plain group (switching scenarios are solved on their
        own structure instead of sharing one), so that adjoint() accepts batches with switching scenarios.
        refine_steps: with factor_precision fp32, refine every Newton-Raphson step in FP64 (the iterates then
        follow the FP64 path); by default only steps whose factorisation perturbed a pivot are refined."""
        self._model = model
        self._handle = Handle()
        self.symmetric = symmetric
        self._opt = model._options(symmetric, error_tolerance, max_iterations, calculation_method, -1,
                                   factor_precision, adjoint_layout=adjoint_layout, refine_steps=refine_steps)
        self._upd = _make_const_dataset(self._handle, update_data, "update", True) if update_data else None
        self._ptr = lib.PGM_amd_batch_create(self._handle.ptr, model._ptr, self._opt.ptr,
                                             None if self._upd is None else self._upd.ptr)
        self._handle.check()
        self.n_scenarios = lib.PGM_amd_batch_n_scenarios(self._ptr)

    def __del__(self):
        if getattr(self, "_ptr", None):
            lib.PGM_amd_batch_destroy(self._ptr)
            self._ptr = None

    def run(self):
        lib.PGM_amd_batch_run(self._handle.ptr, self._ptr)
        self._handle.check()

    def set_output(self, output_component_types=None) -> None:
        """Select what the following runs produce on the device, with the meaning of calculate_power_flow's
        output_component_types: only these components are computed, and components given with attribute
        lists (or as columnar through ComponentAttributeFilterOptions) write only those attributes.
        fetch() then returns that selection."""
        is_batch = self._upd is not None
        out_type, output = self._model._output_arrays(output_component_types, self.symmetric, is_batch,
                                                      self.n_scenarios)
        out_ds = _make_mutable_dataset(self._handle, output, out_type, is_batch, self.n_scenarios)
        lib.PGM_amd_batch_set_output(self._handle.ptr, self._ptr, out_ds.ptr)
        self._handle.check()
        self._output_types = output_component_types

    def fetch(self, output_component_types=None, continue_on_batch_error: bool = True) -> dict:
        is_batch = self._upd is not None
        if output_component_types is None:
            output_component_types = getattr(self, "_output_types", None)
        out_type, output = self._model._output_arrays(output_component_types, self.symmetric, is_batch,
                                                      self.n_scenarios)
        out_ds = _make_mutable_dataset(self._handle, output, out_type, is_batch, self.n_scenarios)
        lib.PGM_amd_batch_fetch(self._handle.ptr, self._ptr, out_ds.ptr)
        code = self._handle.code()
        if code not in (0, 2) or (code == 2 and not continue_on_batch_error):
            raise_regular_error(self._handle.message())
        return output

    def adjoint(self) -> dict:
        """Adjoint gradients of the last run: for every scenario the objective
        L = sum over energised nodes (u_pu - 1)^2 + sum over lines (p_from + p_to) [MW] and its gradients
        by every load_gen's p_specified / q_specified [per W, var] and every line's r1 / x1 [per ohm].
        Returns {component: {attribute: array (n_scenarios, n_component)}}."""
        counts = self._model.all_component_count
        lg = [c for c in ("sym_gen", "asym_gen", "sym_load", "asym_load")]
        n_lg = sum(counts.get(c, 0) for c in lg)
        n_line = counts.get("line", 0)
        S = self.n_scenarios
        gp = np.zeros((S, n_lg))
        gq = np.zeros((S, n_lg))
        gr = np.zeros((S, n_line))
        gx = np.zeros((S, n_line))
        lib.PGM_amd_batch_adjoint(self._handle.ptr, self._ptr, gp.ctypes.data, gq.ctypes.data, gr.ctypes.data,
                                  gx.ctypes.data)
        self._handle.check()
        out = {}
        off = 0
        for c in lg:
            n = counts.get(c, 0)
            if n:
                out[c] = {"p_specified": gp[:, off:off + n], "q_specified": gq[:, off:off + n]}
            off += n
        if n_line:
            out["line"] = {"r1": gr, "x1": gx}
        return out

    def iterations(self) -> np.ndarray:
        out = np.zeros(self.n_scenarios, dtype=np.int64)
        lib.PGM_amd_batch_iterations(self._ptr, out.ctypes.data)
        return out

    def stats(self) -> dict:
        out = np.zeros(14, dtype=np.float64)
        lib.PGM_amd_batch_stats(self._ptr, out.ctypes.data)
        keys = ["kernel_ms", "iterations_sum", "buses", "lu_entries", "groups", "fwd_launch_ms", "fwd_launches",
                "buses_per_scenario", "fwd_bytes", "fwd_total_ms", "fwd_timed", "graph_launches", "refinements",
                "refine_passes"]
        return dict(zip(keys, out.tolist()))


def tap_regulator_ranks(model: PowerGridModel) -> np.ndarray:
    """Host-only ranking of the automatic tap changer: per transformer_tap_regulator (input order) its rank,
    0 next to the sources, -1 when it is off or out of reach (tap_position_optimizer.hpp:209-432)."""
    n = model.all_component_count.get("transformer_tap_regulator", 0)
    ranks = np.zeros(max(n, 1), dtype=np.int64)
    h = model._handle
    lib.PGM_amd_tap_regulator_ranks(h.ptr, model._ptr, ranks.ctypes.data)
    h.check()
    return ranks[:n]


def inspect_topology(model: PowerGridModel, n_node: int):
    """Host-only bus numbering of the product: (node_group, node_pos, fills as (model, a, b))."""
    group = np.zeros(n_node, dtype=np.int64)
    pos = np.zeros(n_node, dtype=np.int64)
    cap = 4 * n_node * n_node + 16
    fills = np.zeros(3 * cap, dtype=np.int64)
    h = model._handle
    nf = lib.PGM_amd_inspect_topology(h.ptr, model._ptr, group.ctypes.data, pos.ctypes.data, fills.ctypes.data, cap)
    h.check()
    return group, pos, [tuple(fills[3 * i:3 * i + 3].tolist()) for i in range(nf)]


def plan_stats(model: PowerGridModel, symmetric: bool = True, target_parallel_tasks: int = 64,
               packed_lanes: int = 0) -> dict:
    """Host-only statistics of the device factorisation schedule of the model's base topology;
    packed_lanes > 0 also checks the interleaved layout of a packed group of that many scenarios."""
    out = np.zeros(7, dtype=np.int64)
    h = model._handle
    lib.PGM_amd_plan_stats(h.ptr, model._ptr, int(symmetric), target_parallel_tasks, packed_lanes,
                           out.ctypes.data)
    h.check()
    keys = ["buses", "lu_entries", "levels", "tasks", "schur_updates", "critical_path", "valid"]
    return dict(zip(keys, out.tolist()))


def minimum_degree_ordering(graph: dict[int, list[int]]):
    keys = np.array(sorted(graph), dtype=np.int64)
    indptr = np.zeros(len(keys) + 1, dtype=np.int64)
    nb = []
    for i, k in enumerate(keys):
        nb.extend(graph[int(k)])
        indptr[i + 1] = len(nb)
    nbs = np.array(nb, dtype=np.int64)
    n_all = len(set(graph) | set(nb))
    alpha = np.zeros(n_all + 1, dtype=np.int64)
    n_alpha = np.zeros(1, dtype=np.int64)
    fills = np.zeros(2 * n_all * n_all + 2, dtype=np.int64)
    nf = lib.PGM_amd_minimum_degree_ordering(len(keys), keys.ctypes.data, indptr.ctypes.data, nbs.ctypes.data,
                                             alpha.ctypes.data, n_alpha.ctypes.data, fills.ctypes.data,
                                             n_all * n_all)
    return alpha[: n_alpha[0]].tolist(), [tuple(fills[2 * i:2 * i + 2].tolist()) for i in range(nf)]


def block_lu_solve(nb: int, row_indptr, col_indices, diag_lu, blocks, rhs, precision: int = 0,
                   no_alloc: bool = False):
    """Device block-sparse LU of a batch of systems sharing one pattern. blocks: (batch, nnz, nb, nb)
    row-major, rhs: (batch, n, nb). Returns (status per system, x). no_alloc: the allocation-free entry point
    (the library's fixed 64 MiB device region; ValueError when the batch does not fit it)."""
    ip = np.ascontiguousarray(row_indptr, dtype=np.int64)
    ci = np.ascontiguousarray(col_indices, dtype=np.int64)
    dg = np.ascontiguousarray(diag_lu, dtype=np.int64)
    blocks = np.asarray(blocks, dtype=np.float64)
    batch = blocks.shape[0]
    data = np.ascontiguousarray(np.transpose(blocks, (0, 1, 3, 2)))  # column-major blocks
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    n = len(ip) - 1
    x = np.zeros((batch, n, nb))
    status = np.zeros(batch, dtype=np.int32)
    fn = lib.PGM_amd_block_lu_solve_noalloc if no_alloc else lib.PGM_amd_block_lu_solve
    rc = fn(nb, n, ip.ctypes.data, ci.ctypes.data, dg.ctypes.data, batch, data.ctypes.data, b.ctypes.data,
            x.ctypes.data, status.ctypes.data, precision)
    if rc == -2:
        raise ValueError("the batch does not fit the device scratch of the allocation-free entry point")
    if rc != 0:
        raise RuntimeError("device block LU failed")
    return status, x
