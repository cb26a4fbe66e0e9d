"""Benchmark: batched Newton-Raphson power flow, scenarios/sec to 1e-8 pu on one MI355X.

A step is one whole job over one batch with the update data already resident in HBM: batch creation
(update readers and id resolution, grouping of the scenarios by topology, topology / ordering / Y-bus /
schedule plans for each group, the device scatter of every scenario's load updates), the linear initial
guess, Newton-Raphson to 1e-8 pu and the output, written as device columns in HBM. The output is the
timing output set, columnar: node u_pu and u_angle, line and transformer p/q at both ends, load p and q.
The update buffers are made resident once, before the timed region (PowerGridModel.make_resident); device
memory of finished batches is reused by the next one. The timed value is the mixed-precision path (FP32
factor and triangular solves, FP64 mismatch and state update); the same run measures:

- the plain FP64 GPU path (fp64_gpu_value), the same job;
- device_resident_value: repeated solves of one created batch (PGM_amd_batch_run), no host work;
- pcie_inclusive: calculate_power_flow from host buffers on a fresh model (upload, whole job, output
  copied back to host), all scenarios, median of several calls after one warm-up;
- the CPU baseline: the FP64 CPU restatement (the oracle), threaded like the reference's batch dispatch
  over all host cores, on a bounded sample; its outputs also give the mixed path's largest error per
  output against the FP64 oracle (mixed_error_vs_oracle) and the FP64 path's iteration-count parity;
- the other workloads (radial_200, meshed_1000, asym_500, feeders_64) with the same legs;
- the automatic tap changer on year_20k_tap and the adjoint on feeders_64.

    python bench.py [--steps K] [--warmup W] [--workload year_20k] [--precision fp32|fp64]
                    [--output timing|records] [--no-tap] [--no-adjoint] [--no-cpu] [--no-configs]
"""


from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "power-grid-model_amd"))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0
# the timing output set: columnar attributes per component (doubles per phase)
TIMING_OUTPUT = {"node": ["u_pu", "u_angle"], "line": ["p_from", "q_from", "p_to", "q_to"],
                 "transformer": ["p_from", "q_from", "p_to", "q_to"], "sym_load": ["p", "q"],
                 "asym_load": ["p", "q"]}
OTHER_WORKLOADS = ("radial_200", "meshed_1000", "asym_500", "feeders_64")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def algorithmic_bytes_per_step(stats, iterations, P, fp32, n_loads, n_scenarios, out_bytes):
    """Sum over scenarios of B_scn = B_io + B_init + k_scn * B_iter (factor written once and read by the
    forward and backward substitutions: 3 E b; voltages and mismatch / increment: 64 P bytes per bus)."""
    E = stats["lu_entries"]
    n = stats["buses_per_scenario"]
    b = (2 * P) ** 2 * (4 if fp32 else 8)
    b_iter = 3 * E * b + 64 * P * n
    b_init = 3 * E * b + 32 * P * n
    b_io = 16 * P * n_loads + out_bytes / max(n_scenarios, 1)
    return float(np.sum(b_io + b_init + iterations * b_iter))


def measured_traffic(workload: str, precision: str):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 --pmc summary of the same
    workload (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy kernel), or None."""
    path = ROOT / "profiles" / f"forward_kernel_{workload}_traffic.md"
    if precision != "fp32" or not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def _leading(update, n):
    """The first n scenarios of an update dataset (uniform arrays or sparse {indptr, data})."""
    out = {}
    for comp, v in update.items():
        if isinstance(v, dict):
            ind = v["indptr"][: n + 1]
            out[comp] = {"indptr": ind.copy(), "data": v["data"][: ind[-1]]}
        else:
            out[comp] = v[:n]
    return out


def _selection(case):
    return {c: a for c, a in TIMING_OUTPUT.items() if c in case.input}


def _out_bytes(case, selection, records):
    P = 1 if case.symmetric else 3
    total = 0
    for comp, arr in case.input.items():
        if not records:
            rec = 8 * P * len(selection.get(comp, []))
        else:
            rec = {"node": 48, "line": 80, "transformer": 80}.get(comp, 48) if case.symmetric else \
                {"node": 128, "line": 208, "transformer": 208}.get(comp, 128)
        total += rec * len(arr) * case.n_scenarios
    return total


def whole_job(model, case, precision, selection, steps, warmup):
    """The timed step: create the batch over the resident update data, select the output, solve. Returns
    the per-step times and the last step's batch (kept for its statistics and outputs)."""
    from power_grid_model_amd import DeviceBatch

    def one():
        t0 = time.perf_counter()
        b = DeviceBatch(model, case.update, symmetric=case.symmetric, factor_precision=precision)
        if selection is not None:
            b.set_output(selection)
        b.run()
        return b, time.perf_counter() - t0

    for _ in range(warmup):
        b, _ = one()
        del b
    times, b = [], None
    for _ in range(steps):
        b = None  # the previous batch hands its device memory back to the model before the next one
        b, t = one()
        times.append(t)
    return np.array(times), b


def device_resident(batch, steps):
    """Repeated solves of one created batch: the solve alone, inputs and structure in HBM."""
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        batch.run()
        times.append(time.perf_counter() - t0)
    return np.array(times)


def pcie_inclusive(case, precision, selection, calls, warmup):
    """calculate_power_flow from host buffers on a model without resident data: upload, the whole job and
    the output columns copied back to the host, all scenarios."""
    from power_grid_model_amd import PowerGridModel

    model = PowerGridModel(case.input)
    times = []
    for i in range(warmup + calls):
        t0 = time.perf_counter()
        model.calculate_power_flow(symmetric=case.symmetric, update_data=case.update, factor_precision=precision,
                                   output_component_types=selection)
        if i >= warmup:
            times.append(time.perf_counter() - t0)
    del model
    return np.array(times)


class ErrorTracker:
    """Largest error of the GPU output against the oracle per output: |u_pu| in pu, angle in rad, flows and
    load powers relative to the largest |value| of their attribute over the compared scenarios."""

    def __init__(self):
        self.abs = {}
        self.scale = {}

    def add(self, gpu, ref, rows):
        for comp, fields in (("node", ("u_pu", "u_angle")), ("line", ("p_from", "q_from", "p_to", "q_to")),
                             ("transformer", ("p_from", "q_from", "p_to", "q_to")), ("sym_load", ("p", "q")),
                             ("asym_load", ("p", "q"))):
            if comp not in gpu or comp not in ref:
                continue
            for f in fields:
                if f not in gpu[comp]:
                    continue
                g = gpu[comp][f][rows]
                r = ref[comp][f]
                d = g - r
                if f == "u_angle":
                    d = np.angle(np.exp(1j * d))
                key = f"{comp}.{f}"
                self.abs[key] = max(self.abs.get(key, 0.0), float(np.max(np.abs(d))) if d.size else 0.0)
                self.scale[key] = max(self.scale.get(key, 0.0), float(np.max(np.abs(r))) if r.size else 0.0)

    def report(self):
        out = {}
        for key, e in self.abs.items():
            out[key] = e if key.startswith("node.") else e / max(self.scale[key], 1e-300)
        return out


def cpu_baseline(case, sample, threads, gpu_out=None, gpu_iters=None, fp64_iters=None):
    """The oracle on `sample` scenarios, `threads` threads, in chunks. With GPU outputs given, the largest
    error per output of the GPU against the oracle, and the iteration counts compared."""
    import oracle_lib as ol

    t_total, done = 0.0, 0
    errors = ErrorTracker()
    it_equal = it_total = extra = 0
    chunk = 256
    for i in range(0, len(sample), chunk):
        part = [int(x) for x in sample[i:i + chunk]]
        upd, k = ol.uniform_update(case.update, case.input, part)
        t0 = time.perf_counter()
        ref, it_ref, st_ref, _ = ol.power_flow(case.input, upd, k, symmetric=case.symmetric, threading=threads)
        t_total += time.perf_counter() - t0
        done += k
        if gpu_out is not None:
            ok = st_ref == 0
            rows = np.array(part)[ok]
            errors.add(gpu_out, {c: {f: ref[c][f][ok] for f in ref[c].dtype.names} for c in ref}, rows)
            if fp64_iters is not None:
                it_equal += int(np.sum(fp64_iters[part] == np.where(ok, it_ref, 0)))
                it_total += k
            if gpu_iters is not None:
                extra += int(np.sum(gpu_iters[rows] - it_ref[ok]))
    res = {"rate": done / t_total, "seconds": t_total, "scenarios": done}
    if gpu_out is not None:
        res["errors"] = errors.report()
        res["fp64_iterations_equal"] = it_equal
        res["fp64_iterations_compared"] = it_total
        res["mixed_extra_iterations"] = extra
    return res


def fetch_columns(batch, selection):
    """Outputs of the batch as {component: {attribute: array (n_scenarios, n)}}."""
    out = batch.fetch(selection)
    res = {}
    for comp, v in out.items():
        if isinstance(v, dict):
            res[comp] = v
        else:
            res[comp] = {f: v[f] for f in v.dtype.names}
    return res


def roofline_of(st, P, fp32):
    fwd_gbs = st["fwd_bytes"] / (st["fwd_total_ms"] * 1e-3) / 1e9 if st["fwd_total_ms"] > 0 else 0.0
    return {"bound": "hbm", "achieved": fwd_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fwd_gbs / HBM_PEAK_GBS,
            "kernel": f"forward_kernel<{P},{'float' if fp32 else 'double'},NR,plain>",
            "bytes_per_launch": st["fwd_bytes"] / max(st["fwd_timed"], 1), "avg_launch_ms": st["fwd_launch_ms"]}


def measure_workload(name, steps, warmup, threads, cpu_seconds, do_cpu):
    """One of the other workloads: mixed and FP64 whole job, the solve alone, the CPU oracle on a bounded
    sample, and the dominant kernel's roofline."""
    from power_grid_model_amd import PowerGridModel, synthetic

    case = synthetic.CONFIGS[name]()
    model = PowerGridModel(case.input)
    model.make_resident(case.update)
    sel = _selection(case)
    P = 1 if case.symmetric else 3
    res = {"workload": name, "scenarios": case.n_scenarios, "symmetric": case.symmetric,
           "load_scale_kappa": case.kappa, "unit": "scenarios/s", "autotune": {}}
    for prec in ("fp32", "fp64"):
        res["autotune"][prec] = model.autotune(case.update, symmetric=case.symmetric, factor_precision=prec)["params"]
        times, b = whole_job(model, case, prec, sel, steps, warmup)
        dr = device_resident(b, steps)
        st = b.stats()
        it = b.iterations()
        key = "mixed" if prec == "fp32" else "fp64"
        res[key] = {"value": case.n_scenarios / float(np.median(times)), "ms_per_step": 1e3 * float(np.median(times)),
                    "device_resident_value": case.n_scenarios / float(np.median(dr)),
                    "mean_iterations": float(np.mean(it[it > 0])) if np.any(it > 0) else 0.0,
                    "pivot_perturbations_refined": int(st.get("refinements", 0)),
                    "failed_scenarios": int(np.sum(it == 0)), "roofline": roofline_of(st, P, prec == "fp32")}
        del b
    res["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    if do_cpu:
        rng = np.random.default_rng(0)
        probe = rng.choice(case.n_scenarios, size=min(16, case.n_scenarios), replace=False)
        r = cpu_baseline(case, probe, threads)
        n_sample = int(min(case.n_scenarios, max(16, r["rate"] * cpu_seconds)))
        sample = rng.choice(case.n_scenarios, size=n_sample, replace=False)
        r = cpu_baseline(case, sample, threads)
        res["cpu_baseline"] = {"value": r["rate"], "unit": "scenarios/s", "cores": threads, "kind": "port",
                               "sample": f"{n_sample} random scenarios ({r['seconds']:.1f} s)"}
        res["gpu_over_cpu"] = res["mixed"]["value"] / r["rate"]
    log(f"[bench] {name}: mixed {res['mixed']['value']:.0f} fp64 {res['fp64']['value']:.0f} scen/s"
        + (f", cpu {res['cpu_baseline']['value']:.0f}" if do_cpu else ""))
    return res


def state_estimation_leg(args, threads):
    """Newton-Raphson state estimation on meshed_1000 (1024 scenarios): sensors placed on the grid, measurements
    from the power flow of each scenario with 0.1% noise; the whole call from host buffers, and the oracle's
    state estimation on a sample of the same scenarios."""
    from power_grid_model_amd import PowerGridModel, synthetic

    se_n = 1024
    pcase = synthetic.meshed_1000(n_scenarios=se_n)
    ptrue = PowerGridModel(pcase.input).calculate_power_flow(
        update_data=pcase.update, output_component_types=["node", "line", "source", "sym_load"])
    scase = synthetic.state_estimation_case(pcase, ptrue)
    smodel = PowerGridModel(scase.input)
    sel = {"node": ["u_pu", "u_angle"], "line": ["p_from", "q_from"]}
    st_times = []
    for k in range(args.steps + 1):
        t0 = time.perf_counter()
        sout = smodel.calculate_state_estimation(update_data=scase.update, calculation_method="newton_raphson",
                                                 output_component_types=sel)
        if k > 0:
            st_times.append(time.perf_counter() - t0)
    t_se = float(np.median(st_times))
    err_u = float(np.max(np.abs(sout["node"]["u_pu"] - ptrue["node"]["u_pu"])))
    state_est = {"workload": scase.name, "scenarios": se_n, "unit": "scenarios/s", "value": se_n / t_se,
                 "seconds": t_se, "step": "calculate_state_estimation from host buffers (newton_raphson, sym)",
                 "sensors": {"voltage": int(len(scase.input["sym_voltage_sensor"])),
                             "power": int(len(scase.input["sym_power_sensor"]))},
                 "max_u_pu_deviation_from_power_flow": err_u}
    log(f"[bench] state estimation {scase.name}: {['%.3f s' % t for t in st_times]}, max |u_pu - pf| {err_u:.2e}")
    if not args.no_cpu:
        # the oracle's state estimation (block sparse gain system, pivot perturbation and refinement as the
        # reference) on the first scenarios, sized to a few seconds; its voltages check the GPU's
        import oracle_lib as ol

        def oracle_se(n):
            upd = {c: v[:n] for c, v in scase.update.items()}
            t0 = time.perf_counter()
            ref, _, st_ref, _ = ol.state_estimation(scase.input, upd, n, threading=threads)
            return ref, st_ref, time.perf_counter() - t0

        _, _, secs = oracle_se(16)
        n_cpu = int(min(se_n, max(16, 16 / secs * args.cpu_seconds / 2)))
        ref, st_ref, secs = oracle_se(n_cpu)
        ok = st_ref == 0
        err_o = float(np.max(np.abs(sout["node"]["u_pu"][:n_cpu][ok] - ref["node"]["u_pu"][ok])))
        state_est["cpu_baseline"] = {"value": n_cpu / secs, "unit": "scenarios/s", "cores": threads, "kind": "port",
                                     "sample": f"first {n_cpu} scenarios of {scase.name} (oracle state "
                                               f"estimation, {threads} threads, {secs:.1f} s)"}
        state_est["max_u_pu_deviation_from_oracle"] = err_o
        log(f"[bench] state estimation cpu baseline {n_cpu / secs:.1f} scen/s on {threads} threads; "
            f"max |u_pu - oracle| {err_o:.2e} over {int(ok.sum())} scenarios")

    return state_est


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="year_20k")
    ap.add_argument("--scenarios", type=int, default=0, help="override the workload's scenario count")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-fp64-leg", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-se", action="store_true", help="skip the state-estimation leg")
    ap.add_argument("--se-only", action="store_true", help="run only the state-estimation leg and print its JSON")
    ap.add_argument("--no-tap", action="store_true", help="skip the automatic tap changer leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-pcie", action="store_true", help="skip the PCIe-inclusive calls")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--output", default="timing", choices=["timing", "records"],
                    help="timing: the columnar timing output set; records: whole records of every component")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default schedule (a profile then sees one launch shape per kernel)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 64 MiB copy kernel (the calibration kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): scenarios shard across GPUs without any "
              "exchange, but this build measures the single-GPU path", file=sys.stderr)
        sys.exit(2)

    from power_grid_model_amd import PowerGridModel, synthetic

    if args.copy_probe:
        from power_grid_model_amd._lib import lib as _clib

        if _clib.PGM_amd_copy_probe(16 << 20) != 0:
            raise RuntimeError("copy probe failed")
        log("[bench] copy probe: copy_probe_kernel read and wrote 67108864 bytes")

    threads = len(os.sched_getaffinity(0))
    if args.se_only:
        print(json.dumps(state_estimation_leg(args, threads)))
        return
    make = synthetic.CONFIGS[args.workload]
    t0 = time.perf_counter()
    case = make(n_scenarios=args.scenarios) if args.scenarios else make()
    log(f"[bench] generated {case.name}: {case.n_scenarios} scenarios in {time.perf_counter() - t0:.1f}s")
    model = PowerGridModel(case.input)
    t0 = time.perf_counter()
    resident = model.make_resident(case.update)
    log(f"[bench] update data made resident: {resident / 1e9:.2f} GB in {time.perf_counter() - t0:.2f}s")
    records = args.output == "records"
    selection = None if records else _selection(case)
    fp32 = args.precision == "fp32"
    P = 1 if case.symmetric else 3

    # schedule parameters picked by measurement for this batch shape (kept by the model for the legs below)
    tuning = {}
    for prec in [] if args.no_autotune else [args.precision] + (["fp64"] if fp32 and not args.no_fp64_leg else []):
        t0 = time.perf_counter()
        tuning[prec] = model.autotune(case.update, symmetric=case.symmetric, factor_precision=prec)
        tuning[prec]["seconds"] = time.perf_counter() - t0
        log(f"[bench] autotune {prec}: {tuning[prec]['params']} in {tuning[prec]['seconds']:.1f}s; trials "
            f"{[(tuple(p.values()), round(ms, 2)) for p, ms in tuning[prec]['trials']]}")

    # the headline: the whole job, inputs resident in HBM
    times, batch = whole_job(model, case, args.precision, selection, args.steps, args.warmup)
    t_step = float(np.median(times))
    value = case.n_scenarios / t_step
    log(f"[bench] {args.precision} whole job: {['%.1f ms' % (1e3 * t) for t in times]}")
    # the solve alone on the last batch (roofline of the dominant kernel from its HIP events), with the
    # Newton-Raphson iterations as graph launches (default) and as one launch per level
    dr = device_resident(batch, args.steps)
    st = batch.stats()
    it = batch.iterations()
    log(f"[bench] {args.precision} solve only: {['%.1f ms' % (1e3 * t) for t in dr]} stats {st}")
    os.environ["PGM_AMD_GRAPH"] = "0"
    dr_levels = device_resident(batch, args.steps)
    # the dominant kernel's duration from the events around its direct launches on its stream: events
    # captured as graph nodes also time the node hand-over around the kernel (about 0.2 ms per launch)
    graph_launch_ms = st["fwd_launch_ms"]
    st = batch.stats()
    del os.environ["PGM_AMD_GRAPH"]
    log(f"[bench] {args.precision} solve only, one launch per level: {['%.1f ms' % (1e3 * t) for t in dr_levels]}")
    failed = int(np.sum(it == 0))
    n_loads = sum(len(case.input.get(c, [])) for c in ("sym_load", "asym_load", "sym_gen", "asym_gen"))
    step_bytes = algorithmic_bytes_per_step(st, it.astype(np.float64), P, fp32, n_loads, case.n_scenarios,
                                            _out_bytes(case, selection or {}, records))
    roof = roofline_of(st, P, fp32)
    roof["traffic"] = measured_traffic(case.name, args.precision)
    roof["graph_event_launch_ms"] = graph_launch_ms
    mixed_out = fetch_columns(batch, selection) if (not args.no_cpu and not records) else None
    mixed_iters = it.copy()
    del batch

    fp64_value = fp64_dr = None
    fp64_iters = None
    if fp32 and not args.no_fp64_leg:
        t64, b64 = whole_job(model, case, "fp64", selection, args.steps, args.warmup)
        fp64_value = case.n_scenarios / float(np.median(t64))
        d64 = device_resident(b64, args.steps)
        fp64_dr = case.n_scenarios / float(np.median(d64))
        fp64_iters = b64.iterations().copy()
        log(f"[bench] fp64 whole job: {['%.1f ms' % (1e3 * t) for t in t64]}; solve only "
            f"{['%.1f ms' % (1e3 * t) for t in d64]}")
        del b64

    # PCIe-inclusive: calculate_power_flow from host buffers, all scenarios
    pcie = None
    if not args.no_pcie:
        comps = selection if selection is not None else \
            [c for c in ("node", "line", "transformer", "sym_load", "asym_load") if c in case.input]
        pcie = {"unit": "scenarios/s", "scenarios": case.n_scenarios, "outputs": comps}
        for prec in ([args.precision] + (["fp64"] if fp32 and not args.no_fp64_leg else [])):
            tp = pcie_inclusive(case, prec, comps, args.steps, 1)
            key = "value" if prec == args.precision else "fp64_value"
            pcie[key] = case.n_scenarios / float(np.median(tp))
            pcie[("seconds" if prec == args.precision else "fp64_seconds")] = float(np.median(tp))
            log(f"[bench] calculate_power_flow from host buffers ({prec}): "
                f"{['%.3f s' % t for t in tp]}")

    state_est = None if args.no_se else state_estimation_leg(args, threads)

    # adjoint gradients on the uneven-feeder batch (64 feeders, 128 scenarios): the solve, then one
    # adjoint pass (re-factorisation in FP64 at the converged state and the transposed sweeps)
    adjoint = None
    if not args.no_adjoint:
        from power_grid_model_amd import DeviceBatch

        fcase = synthetic.feeders_64()
        fmodel = PowerGridModel(fcase.input)
        fb = DeviceBatch(fmodel, fcase.update, factor_precision=args.precision)
        fb.set_output({"node": ["u_pu"], "line": ["p_from", "p_to"]})
        fb.run()
        solve, adj = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            fb.run()
            solve.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            g = fb.adjoint()
            adj.append(time.perf_counter() - t0)
        n_par = sum(a.size for comp in g.values() for a in comp.values())
        adjoint = {"workload": fcase.name, "scenarios": fcase.n_scenarios, "solve_ms": 1e3 * float(np.median(solve)),
                   "adjoint_ms": 1e3 * float(np.median(adj)), "gradient_entries": int(n_par)}
        log(f"[bench] adjoint on {fcase.name}: solve {adjoint['solve_ms']:.1f} ms, adjoint {adjoint['adjoint_ms']:.1f} ms")
        del fb, fmodel

    # automatic tap changer on year_20k_tap (the 110/10.5 kV transformer regulated): calculate_power_flow
    # from host buffers with a tap changing strategy; every round is one device batch over the scenarios
    # still searching. The oracle's tap changer runs beside it on a bounded sample.
    tap = None
    if not args.no_tap:
        tcase = synthetic.year_20k_tap()
        tmodel = PowerGridModel(tcase.input)
        tsel = dict(_selection(tcase), transformer_tap_regulator=["tap_pos"])
        tap = {"workload": tcase.name, "scenarios": tcase.n_scenarios, "unit": "scenarios/s", "strategies": {}}
        for strategy in ("any_valid_tap", "min_voltage_tap"):
            t0 = time.perf_counter()
            tout = tmodel.calculate_power_flow(update_data=tcase.update, tap_changing_strategy=strategy,
                                               factor_precision="fp64", output_component_types=tsel)
            secs = time.perf_counter() - t0
            taps = tout["transformer_tap_regulator"]["tap_pos"][:, 0]
            tap["strategies"][strategy] = {"value": tcase.n_scenarios / secs, "seconds": secs,
                                           "distinct_taps": int(len(np.unique(taps)))}
            log(f"[bench] tap changer {strategy}: {tcase.n_scenarios} scenarios in {secs:.2f}s")
        if not args.no_cpu:
            import oracle_lib as ol
            n_tap = 256
            upd_u, k = ol.uniform_update(tcase.update, tcase.input, list(range(n_tap)))
            t0 = time.perf_counter()
            ol.power_flow_tap(tcase.input, upd_u, k, strategy="any_valid_tap", threading=threads)
            secs = time.perf_counter() - t0
            tap["cpu_baseline"] = {"value": k / secs, "unit": "scenarios/s", "cores": threads, "kind": "port",
                                   "sample": f"first {k} scenarios of {tcase.name}, any_valid_tap (oracle tap "
                                             f"changer, {threads} threads, {secs:.1f} s)"}
            log(f"[bench] tap changer cpu baseline {k / secs:.1f} scen/s on {threads} threads")
        del tmodel

    # CPU baseline on a bounded sample of the same workload; its outputs pin the mixed path's error
    cpu = None
    cmp = None
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        probe = rng.choice(case.n_scenarios, size=min(32, case.n_scenarios), replace=False)
        rate = cpu_baseline(case, probe, threads)["rate"]
        n_sample = int(min(case.n_scenarios, max(32, rate * args.cpu_seconds)))
        sample = rng.choice(case.n_scenarios, size=n_sample, replace=False)
        cmp = cpu_baseline(case, sample, threads, gpu_out=mixed_out, gpu_iters=mixed_iters if fp32 else None,
                           fp64_iters=fp64_iters if fp32 else mixed_iters)
        cpu = {"value": cmp["rate"], "unit": "scenarios/s", "cores": threads, "cpu": cpu_model(), "kind": "port",
               "sample": f"{n_sample} random scenarios of {case.name} (FP64 CPU restatement, {threads} threads, "
                         f"{cmp['seconds']:.1f} s)"}
        log(f"[bench] cpu baseline {cmp['rate']:.1f} scen/s on {threads} threads ({n_sample} scenarios, "
            f"{cmp['seconds']:.1f}s); errors {cmp.get('errors')}")
        del mixed_out

    configs = None
    if not args.no_configs and args.workload == "year_20k":
        configs = {}
        for name in OTHER_WORKLOADS:
            configs[name] = measure_workload(name, args.steps, args.warmup, threads, 3.0, not args.no_cpu)

    line = {
        "metric": "NR power-flow scenarios/sec at 1e-8 pu",
        "value": value,
        "unit": "scenarios/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor / f64 mismatch" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": case.name, "scenarios": case.n_scenarios,
                   "buses": int(st["buses_per_scenario"]), "lu_entries": int(st["lu_entries"]),
                   "symmetric": case.symmetric, "load_scale_kappa": case.kappa, "topology_groups": int(st["groups"]),
                   "err_tol": 1e-8, "failed_scenarios": failed, "mean_iterations": float(np.mean(it[it > 0])),
                   "pivot_perturbations_refined": int(st.get("refinements", 0)),
                   "step": "whole job with the update data resident in HBM: batch creation (host grouping, "
                           "topology and plans; device scatter of the load updates), solve, output columns in HBM",
                   "output": ("columnar " + ", ".join(f"{c}{{{','.join(a)}}}" for c, a in selection.items()))
                   if selection is not None else "whole records of every component"},
        "device_resident_value": case.n_scenarios / float(np.median(dr)),
        "device_resident_level_launches_value": case.n_scenarios / float(np.median(dr_levels)),
        "fp64_gpu_value": fp64_value,
        "fp64_device_resident_value": fp64_dr,
        "mixed_over_fp64": (value / fp64_value) if fp64_value else None,
        "mixed_error_vs_oracle": cmp.get("errors") if cmp else None,
        "iteration_parity": ({"fp64_equal": cmp["fp64_iterations_equal"], "compared": cmp["fp64_iterations_compared"],
                              "mixed_extra_iterations": cmp["mixed_extra_iterations"]} if cmp else None),
        "autotune": tuning,
        "roofline": roof,
        "step_roofline": {"algorithmic_bytes": step_bytes, "achieved_gbs": step_bytes / t_step / 1e9,
                          "frac": step_bytes / t_step / 1e9 / HBM_PEAK_GBS},
        "pcie_inclusive": pcie,
        "adjoint": adjoint,
        "state_estimation": state_est,
        "tap_changer": tap,
        "configs": configs,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "pcie_gpu_over_cpu": (pcie["value"] / cpu["value"]) if (cpu and pcie) else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
