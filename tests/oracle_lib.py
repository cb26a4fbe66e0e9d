"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module: the oracle is the
checker and the CPU baseline, never part of the product path.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from pgm_formats import ASYM_OUTPUT, COMPONENT_ORDER, INPUT, SYM_OUTPUT, UPDATE

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "_build" / "liboracle.so"

N_COMP = len(COMPONENT_ORDER)
_i64 = C.c_int64
_p = C.c_void_p


class _Const(C.Structure):
    _fields_ = [("data", _p * N_COMP), ("count", _i64 * N_COMP)]


class _Mut(C.Structure):
    _fields_ = [("data", _p * N_COMP), ("count", _i64 * N_COMP)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        _lib = C.CDLL(str(LIB_PATH))
        _lib.or_power_flow.restype = C.c_int
        _lib.or_power_flow.argtypes = [C.POINTER(_Const), C.c_double, C.POINTER(_Const), _i64, C.c_int,
                                       C.c_double, _i64, _i64, C.POINTER(_Mut), _p, _p, _p, _i64]
        _lib.or_power_flow_method.restype = C.c_int
        _lib.or_power_flow_method.argtypes = [C.POINTER(_Const), C.c_double, C.POINTER(_Const), _i64, C.c_int,
                                              C.c_double, _i64, _i64, C.c_int, C.POINTER(_Mut), _p, _p, _p, _i64]
        _lib.or_power_flow_tap.restype = C.c_int
        _lib.or_power_flow_tap.argtypes = [C.POINTER(_Const), C.c_double, C.POINTER(_Const), _i64, C.c_int,
                                           C.c_double, _i64, _i64, C.c_int, C.c_int, _p, _i64, _p, _i64,
                                           C.POINTER(_Mut), _p, _p, _p, _p, _i64]
        _lib.or_state_estimation.restype = C.c_int
        _lib.or_state_estimation.argtypes = [C.POINTER(_Const), C.c_double, C.POINTER(_Const), _i64, C.c_double,
                                             _i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p, _i64, C.POINTER(_Mut), _p, _p,
                                             _p, _p, _p, _i64]
        _lib.or_set_se_dense.restype = None
        _lib.or_set_se_dense.argtypes = [C.c_int]
        _lib.or_se_counters.restype = None
        _lib.or_se_counters.argtypes = [_p, _p, C.c_int]
        _lib.or_minimum_degree_ordering.restype = _i64
        _lib.or_minimum_degree_ordering.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _i64]
        _lib.or_build_topology.restype = _i64
        _lib.or_build_topology.argtypes = [_i64, _i64, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _i64]
        _lib.or_ybus_structure.restype = _i64
        _lib.or_ybus_structure.argtypes = [_i64, _i64, _p, _i64, _p, _i64, _p, _i64] + [_p] * 10
        _lib.or_block_lu_solve.restype = C.c_int
        _lib.or_block_lu_solve.argtypes = [_i64, _i64, _p, _p, _p, _p, _p, _p]
        _lib.or_nr_math_model.restype = _i64
        _lib.or_nr_math_model.argtypes = [C.c_int, _i64, _i64, _p, _p, _i64, _p, _p, _i64, _p, _p, _p, _i64, _p,
                                          _p, _p, _i64, _p, C.c_double, _i64, _p, _p, _p, _p, _p, _p]
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


METHODS = {"newton_raphson": 0, "linear": 1, "iterative_current": 2, "linear_current": 3}


def power_flow(input_data: dict, update: dict | None = None, n_scenarios: int = 1, symmetric: bool = True,
               err_tol: float = 1e-8, max_iter: int = 20, threading: int = -1, system_frequency: float = 50.0,
               method: str = "newton_raphson"):
    """Run the oracle. update: {component: array of shape (n_scenarios, rows_per_scenario)}.

    Returns (output dict {component: array (n_scenarios, n)}, iterations, status, messages)."""
    keep, cin, cup, cout, out, n_scenarios = _datasets(input_data, update, n_scenarios, symmetric)
    iters = np.zeros(n_scenarios, dtype=np.int64)
    status = np.zeros(n_scenarios, dtype=np.int32)
    stride = 512
    msgs = np.zeros(n_scenarios * stride, dtype=np.uint8)
    rc = lib().or_power_flow_method(C.byref(cin), system_frequency, C.byref(cup) if cup is not None else None,
                                    n_scenarios, int(symmetric), err_tol, max_iter, threading, METHODS[method],
                                    C.byref(cout), _ptr(iters), _ptr(status), _ptr(msgs), stride)
    messages = [bytes(msgs[i * stride:(i + 1) * stride]).split(b"\0")[0].decode() for i in range(n_scenarios)]
    if rc != 0:
        raise RuntimeError(messages[0])
    return out, iters, status, messages


TAP_STRATEGIES = {"any_valid_tap": 1, "min_voltage_tap": 2, "max_voltage_tap": 3, "fast_any_tap": 4}


def power_flow_tap(input_data: dict, update: dict | None = None, n_scenarios: int = 1, symmetric: bool = True,
                   strategy: str = "any_valid_tap", err_tol: float = 1e-8, max_iter: int = 20, threading: int = -1,
                   system_frequency: float = 50.0, method: str = "newton_raphson"):
    """Power flow with automatic tap changing. input_data / update may hold transformer_tap_regulator.

    Returns (output dict, tap positions (n_scenarios, n_regulator), iterations, status, messages); status
    4 marks a scenario whose optimizer failed (MaxIterationReached, invalid regulator configuration)."""
    keep, cin, cup, cout, out, n_scenarios = _datasets(input_data, update, n_scenarios, symmetric)
    reg = input_data.get("transformer_tap_regulator")
    reg = np.ascontiguousarray(reg if reg is not None else np.zeros(0, INPUT["transformer_tap_regulator"]),
                               dtype=INPUT["transformer_tap_regulator"])
    ru = None if update is None else update.get("transformer_tap_regulator")
    if ru is not None:
        ru = np.ascontiguousarray(ru, dtype=UPDATE["transformer_tap_regulator"])
    taps = np.zeros((n_scenarios, max(len(reg), 1)), dtype=np.int8)
    iters = np.zeros(n_scenarios, dtype=np.int64)
    status = np.zeros(n_scenarios, dtype=np.int32)
    stride = 512
    msgs = np.zeros(n_scenarios * stride, dtype=np.uint8)
    rc = lib().or_power_flow_tap(C.byref(cin), system_frequency, C.byref(cup) if cup is not None else None,
                                 n_scenarios, int(symmetric), err_tol, max_iter, threading, METHODS[method],
                                 TAP_STRATEGIES[strategy], _ptr(reg), len(reg), _ptr(ru),
                                 0 if ru is None else ru.shape[1], C.byref(cout), _ptr(taps), _ptr(iters),
                                 _ptr(status), _ptr(msgs), stride)
    messages = [bytes(msgs[i * stride:(i + 1) * stride]).split(b"\0")[0].decode() for i in range(n_scenarios)]
    if rc != 0:
        raise RuntimeError(messages[0])
    return out, taps[:, :len(reg)], iters, status, messages


def state_estimation(input_data: dict, update: dict | None = None, n_scenarios: int = 1, err_tol: float = 1e-8,
                     max_iter: int = 20, system_frequency: float = 50.0, threading: int = -1):
    """Symmetric Newton-Raphson state estimation on the oracle. input_data / update may hold sym_voltage_sensor
    and sym_power_sensor (update arrays of shape (n_scenarios, rows)).

    Returns (output dict with the sensors' residuals, iterations, status, messages); status 1 marks a
    SparseMatrixError, 2 a diverged iteration, 3 any other failure (NotObservableError and the like)."""
    keep, cin, cup, cout, out, n_scenarios = _datasets(input_data, update, n_scenarios, True)

    def sensor(name, table):
        arr = input_data.get(name)
        return np.ascontiguousarray(arr if arr is not None else np.zeros(0, table[name]), dtype=table[name])

    def sensor_update(name):
        arr = None if update is None else update.get(name)
        return None if arr is None else np.ascontiguousarray(arr, dtype=UPDATE[name])

    vs, ps = sensor("sym_voltage_sensor", INPUT), sensor("sym_power_sensor", INPUT)
    vu, pu = sensor_update("sym_voltage_sensor"), sensor_update("sym_power_sensor")
    vs_out = np.zeros((n_scenarios, len(vs)), dtype=SYM_OUTPUT["sym_voltage_sensor"])
    ps_out = np.zeros((n_scenarios, len(ps)), dtype=SYM_OUTPUT["sym_power_sensor"])
    iters = np.zeros(n_scenarios, dtype=np.int64)
    status = np.zeros(n_scenarios, dtype=np.int32)
    stride = 512
    msgs = np.zeros(n_scenarios * stride, dtype=np.uint8)
    rc = lib().or_state_estimation(C.byref(cin), system_frequency, C.byref(cup) if cup is not None else None,
                                   n_scenarios, err_tol, max_iter, threading, _ptr(vs), len(vs), _ptr(vu),
                                   0 if vu is None else vu.shape[1], _ptr(ps), len(ps), _ptr(pu),
                                   0 if pu is None else pu.shape[1], C.byref(cout), _ptr(vs_out), _ptr(ps_out),
                                   _ptr(iters), _ptr(status), _ptr(msgs), stride)
    messages = [bytes(msgs[i * stride:(i + 1) * stride]).split(b"\0")[0].decode() for i in range(n_scenarios)]
    if rc != 0:
        raise RuntimeError(messages[0])
    if len(vs):
        out["sym_voltage_sensor"] = vs_out
    if len(ps):
        out["sym_power_sensor"] = ps_out
    return out, iters, status, messages


def set_se_dense(dense: bool) -> None:
    """Solve the state estimation's gain system densely (a second restatement) instead of with the block sparse
    LU, pivot perturbation and refinement."""
    lib().or_set_se_dense(int(dense))


def se_counters(reset: bool = True) -> tuple[int, int]:
    """(sparse gain factorisations that perturbed a pivot, refinement solves) since the last reset."""
    a = C.c_longlong(0)
    b = C.c_longlong(0)
    lib().or_se_counters(C.addressof(a), C.addressof(b), int(reset))
    return a.value, b.value


def _datasets(input_data: dict, update: dict | None, n_scenarios: int, symmetric: bool):
    keep = []
    cin = _Const()
    for k, comp in enumerate(COMPONENT_ORDER):
        arr = input_data.get(comp)
        if arr is not None and len(arr):
            arr = np.ascontiguousarray(arr, dtype=INPUT[comp])
            keep.append(arr)
            cin.data[k] = arr.ctypes.data
            cin.count[k] = len(arr)
    cup = None
    if update is not None:
        cup = _Const()
        for k, comp in enumerate(COMPONENT_ORDER):
            arr = update.get(comp)
            if arr is not None and arr.size:
                arr = np.ascontiguousarray(arr, dtype=UPDATE[comp])
                keep.append(arr)
                cup.data[k] = arr.ctypes.data
                cup.count[k] = arr.shape[1]
    else:
        n_scenarios = 1
    odt = SYM_OUTPUT if symmetric else ASYM_OUTPUT
    cout = _Mut()
    out = {}
    for k, comp in enumerate(COMPONENT_ORDER):
        arr = input_data.get(comp)
        if arr is not None and len(arr):
            o = np.zeros((n_scenarios, len(arr)), dtype=odt[comp])
            out[comp] = o
            cout.data[k] = o.ctypes.data
            cout.count[k] = len(arr)
    return keep, cin, cup, cout, out, n_scenarios


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
    fills = np.zeros(2 * n_all * n_all, dtype=np.int64)
    nf = lib().or_minimum_degree_ordering(len(keys), _ptr(keys), _ptr(indptr), _ptr(nbs), _ptr(alpha),
                                          _ptr(n_alpha), _ptr(fills), n_all * n_all)
    return alpha[: n_alpha[0]].tolist(), [tuple(fills[2 * i:2 * i + 2].tolist()) for i in range(nf)]


def build_topology(n_node, branches, sources, phase_shift=None, statuses=None, source_status=None):
    br = np.array(branches, dtype=np.int64).reshape(-1, 2)
    nb = len(br)
    st = np.ones((nb, 2), dtype=np.int8) if statuses is None else np.array(statuses, dtype=np.int8)
    ps = np.zeros(nb) if phase_shift is None else np.array(phase_shift, dtype=np.float64)
    src = np.array(sources, dtype=np.int64)
    sst = np.ones(len(src), dtype=np.int8) if source_status is None else np.array(source_status, dtype=np.int8)
    f = np.ascontiguousarray(br[:, 0])
    t = np.ascontiguousarray(br[:, 1])
    fs = np.ascontiguousarray(st[:, 0])
    ts = np.ascontiguousarray(st[:, 1])
    group = np.zeros(n_node, dtype=np.int64)
    pos = np.zeros(n_node, dtype=np.int64)
    nmm = np.zeros(1, dtype=np.int64)
    cap = 4 * n_node * n_node + 16
    fill = np.zeros(3 * cap, dtype=np.int64)
    nf = lib().or_build_topology(n_node, nb, _ptr(f), _ptr(t), _ptr(fs), _ptr(ts), _ptr(ps), len(src), _ptr(src),
                                 _ptr(sst), _ptr(group), _ptr(pos), _ptr(nmm), _ptr(fill), cap)
    fills = [tuple(fill[3 * i:3 * i + 3].tolist()) for i in range(nf)]
    return group.tolist(), pos.tolist(), int(nmm[0]), fills


def ybus_structure(n_bus, branch_bus, shunt_bus, fill_in):
    bb = np.array(branch_bus, dtype=np.int64).reshape(-1)
    sb = np.array(shunt_bus, dtype=np.int64)
    fi = np.array(fill_in, dtype=np.int64).reshape(-1)
    cap = 4 * len(bb) + 2 * len(fi) + n_bus + len(sb) + 4
    outs = {k: np.zeros(cap, dtype=np.int64) for k in ("row_indptr_lu", "col_indices_lu", "diag_lu", "map_lu_y_bus",
                                                       "lu_transpose_entry", "row_indptr", "col_indices",
                                                       "bus_entry", "y_bus_entry_indptr")}
    nnz = np.zeros(1, dtype=np.int64)
    nnz_lu = lib().or_ybus_structure(n_bus, len(bb) // 2, _ptr(bb), len(sb), _ptr(sb), len(fi) // 2, _ptr(fi), cap,
                                     *[_ptr(outs[k]) for k in ("row_indptr_lu", "col_indices_lu", "diag_lu",
                                                               "map_lu_y_bus", "lu_transpose_entry", "row_indptr",
                                                               "col_indices", "bus_entry", "y_bus_entry_indptr")],
                                     _ptr(nnz))
    n = int(nnz[0])
    return {
        "row_indptr_lu": outs["row_indptr_lu"][: n_bus + 1].tolist(),
        "col_indices_lu": outs["col_indices_lu"][:nnz_lu].tolist(),
        "diag_lu": outs["diag_lu"][:n_bus].tolist(),
        "map_lu_y_bus": outs["map_lu_y_bus"][:nnz_lu].tolist(),
        "lu_transpose_entry": outs["lu_transpose_entry"][:nnz_lu].tolist(),
        "row_indptr": outs["row_indptr"][: n_bus + 1].tolist(),
        "col_indices": outs["col_indices"][:n].tolist(),
        "bus_entry": outs["bus_entry"][:n_bus].tolist(),
        "y_bus_entry_indptr": outs["y_bus_entry_indptr"][: n + 1].tolist(),
    }


def block_lu_solve(nb, row_indptr, col_indices, diag_lu, blocks, rhs):
    """blocks: (nnz, nb, nb) row-major numpy -> converted to column-major; rhs (n, nb)."""
    n = len(row_indptr) - 1
    ip = np.array(row_indptr, dtype=np.int64)
    ci = np.array(col_indices, dtype=np.int64)
    dg = np.array(diag_lu, dtype=np.int64)
    data = np.ascontiguousarray(np.transpose(np.asarray(blocks, dtype=np.float64), (0, 2, 1)))
    b = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64))
    x = np.zeros((n, nb))
    rc = lib().or_block_lu_solve(nb, n, _ptr(ip), _ptr(ci), _ptr(dg), _ptr(data), _ptr(b), _ptr(x))
    return rc, x


def nr_math_model(p, n_bus, branch_bus, branch_param, shunt_bus, shunt_param, source_bus, source_y, u_source,
                  load_bus, load_type, s_inj, fill_in=(), err_tol=1e-12, max_iter=20):
    """Complex arrays: branch_param (nbr,4,p,p), shunt_param (nsh,p,p), source_y (nsrc,2), u_source (nsrc,),
    s_inj (nload,p). Returns (iterations or negative error, results dict)."""
    def c2r(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        return np.ascontiguousarray(a.view(np.float64))
    bb = np.array(branch_bus, dtype=np.int64).reshape(-1)
    bp = c2r(branch_param)
    sb = np.array(shunt_bus, dtype=np.int64)
    sp = c2r(shunt_param) if len(shunt_bus) else np.zeros(2)
    srcb = np.array(source_bus, dtype=np.int64)
    sy = c2r(source_y)
    us = c2r(u_source)
    lb = np.array(load_bus, dtype=np.int64)
    lt = np.array(load_type, dtype=np.int8)
    si = c2r(s_inj)
    fi = np.array(fill_in, dtype=np.int64).reshape(-1)
    nbr, nsh, nl, nsrc = len(bb) // 2, len(sb), len(lb), len(srcb)
    u = np.zeros((n_bus, p), dtype=np.complex128)
    bf = np.zeros((nbr, 4, p), dtype=np.complex128)
    sf = np.zeros((max(nsh, 1), 2, p), dtype=np.complex128)
    lf = np.zeros((max(nl, 1), 2, p), dtype=np.complex128)
    srf = np.zeros((max(nsrc, 1), 2, p), dtype=np.complex128)
    inj = np.zeros((n_bus, p), dtype=np.complex128)
    it = lib().or_nr_math_model(p, n_bus, nbr, _ptr(bb), _ptr(bp), nsh, _ptr(sb), _ptr(sp), nsrc, _ptr(srcb),
                                _ptr(sy), _ptr(us), nl, _ptr(lb), _ptr(lt), _ptr(si), len(fi) // 2, _ptr(fi),
                                err_tol, max_iter, _ptr(u), _ptr(bf), _ptr(sf), _ptr(lf), _ptr(srf), _ptr(inj))
    return it, {"u": u, "branch": bf, "shunt": sf[:nsh], "load": lf[:nl], "source": srf[:nsrc],
                "bus_injection": inj}


def uniform_update(update: dict, input_data: dict, scenarios=None) -> tuple[dict, int]:
    """Convert a batch update in any of the package's forms (uniform 2D arrays or sparse {"indptr", "data"})
    into the oracle's uniform form, padding scenarios with no-op rows (ids of untouched components,
    null attributes). Optionally keep only the listed scenarios."""
    from pgm_formats import batch_to_uniform

    n = None
    for v in update.values():
        n = (len(v["indptr"]) - 1) if isinstance(v, dict) else v.shape[0]
        break
    sel = list(range(n)) if scenarios is None else list(scenarios)
    per_scn = [dict() for _ in sel]
    for comp, v in update.items():
        for k, s in enumerate(sel):
            if isinstance(v, dict):
                per_scn[k][comp] = v["data"][v["indptr"][s]:v["indptr"][s + 1]]
            else:
                per_scn[k][comp] = v[s]
    return batch_to_uniform(per_scn, input_data), len(sel)
