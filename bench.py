"""Benchmark: batched Newton-Raphson power flow, scenarios/sec to 1e-8 pu on one MI355X.

A step is one pass of the hot path over one batch: every scenario of the workload solved to
convergence (linear initial guess, NR iterations, output) with the inputs already resident in HBM
(PGM_amd_batch_run). The output of a step is the timing output set, columnar: node u_pu and u_angle,
line p/q at both ends, load p and q (--output records: every component as whole records, also
reported as records_value). The timed value is the mixed-precision path (FP32 factor and triangular
solves, FP64 mismatch and state update); the plain FP64 GPU path and the CPU baseline (the FP64 CPU
restatement, threaded like the reference's batch dispatch over all host cores, on a bounded sample)
are measured in the same run.

    python bench.py [--steps K] [--warmup W] [--workload year_20k] [--precision fp32|fp64]
                    [--output timing|records] [--no-tap] [--no-adjoint] [--no-cpu]

The line also carries the automatic tap changer on year_20k_tap (scenarios/s of calculate_power_flow with
a tap changing strategy, from host buffers), the adjoint on feeders_64 and the PCIe-inclusive rate.
"""


from __future__ import annotations

import argparse
import json
import os
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def cpu_baseline(case, sample_scenarios, threads):
    import oracle_lib as ol

    sel = list(sample_scenarios)
    t_total = 0.0
    done = 0
    chunk = 256
    for i in range(0, len(sel), chunk):
        part = sel[i:i + chunk]
        upd, k = ol.uniform_update(case.update, case.input, part)
        t0 = time.perf_counter()
        ol.power_flow(case.input, upd, k, symmetric=case.symmetric, threading=threads)
        t_total += time.perf_counter() - t0
        done += k
    return done / t_total, t_total


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
    ap.add_argument("--no-tap", action="store_true", help="skip the automatic tap changer leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--e2e-fraction", type=float, default=0.2,
                    help="share of the scenarios timed through calculate_power_flow from host buffers (0: skip)")
    ap.add_argument("--output", default="timing", choices=["timing", "records"],
                    help="timing: the columnar timing output set; records: whole records of every component")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 64 MiB copy kernel (the calibration kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): scenarios shard across GPUs without any "
              "exchange, but this build measures the single-GPU path", file=sys.stderr)
        sys.exit(2)

    from power_grid_model_amd import DeviceBatch, PowerGridModel, synthetic

    if args.copy_probe:
        from power_grid_model_amd._lib import lib as _clib

        if _clib.PGM_amd_copy_probe(16 << 20) != 0:
            raise RuntimeError("copy probe failed")
        log("[bench] copy probe: copy_probe_kernel read and wrote 67108864 bytes")

    make = synthetic.CONFIGS[args.workload]
    t0 = time.perf_counter()
    case = make(n_scenarios=args.scenarios) if args.scenarios else make()
    log(f"[bench] generated {case.name}: {case.n_scenarios} scenarios in {time.perf_counter() - t0:.1f}s")
    model = PowerGridModel(case.input)

    selection = {c: a for c, a in TIMING_OUTPUT.items() if c in case.input}

    def measure(precision, output=args.output, steps=args.steps):
        t0 = time.perf_counter()
        batch = DeviceBatch(model, case.update, symmetric=case.symmetric, factor_precision=precision)
        if output == "timing":
            batch.set_output(selection)
        log(f"[bench] {precision}: batch created in {time.perf_counter() - t0:.1f}s (output: {output})")
        for _ in range(args.warmup):
            batch.run()
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            batch.run()
            times.append(time.perf_counter() - t0)
        st = batch.stats()
        it = batch.iterations()
        log(f"[bench] {precision}: steps {['%.1f ms' % (1e3 * t) for t in times]} stats {st}")
        return batch, np.array(times), st, it

    batch, times, st, it = measure(args.precision)
    t_step = float(np.median(times))
    value = case.n_scenarios / t_step
    failed = int(np.sum(it == 0))
    fp32 = args.precision == "fp32"
    P = 1 if case.symmetric else 3
    n_loads = sum(len(case.input.get(c, [])) for c in ("sym_load", "asym_load", "sym_gen", "asym_gen"))
    out_bytes = 0
    for comp, arr in case.input.items():
        if args.output == "timing":
            rec = 8 * P * len(selection.get(comp, []))
        else:
            rec = {"node": 48, "line": 80, "transformer": 80}.get(comp, 48) if case.symmetric else \
                {"node": 128, "line": 208, "transformer": 208}.get(comp, 128)
        out_bytes += rec * len(arr) * case.n_scenarios
    step_bytes = algorithmic_bytes_per_step(st, it.astype(np.float64), P, fp32, n_loads, case.n_scenarios,
                                            out_bytes)
    fwd_gbs = st["fwd_bytes"] / (st["fwd_total_ms"] * 1e-3) / 1e9 if st["fwd_total_ms"] > 0 else 0.0
    fwd_bytes_per_launch = st["fwd_bytes"] / max(st["fwd_timed"], 1)
    del batch

    # the same path writing every component as whole records (3.1 MB per year_20k scenario)
    records_value = None
    if args.output == "timing":
        brec, trec, _, _ = measure(args.precision, output="records", steps=max(1, min(args.steps, 5)))
        records_value = case.n_scenarios / float(np.median(trec))
        del brec

    fp64_value = None
    if fp32 and not args.no_fp64_leg:
        b64, t64, st64, it64 = measure("fp64")
        fp64_value = case.n_scenarios / float(np.median(t64))
        del b64

    # PCIe-inclusive rate: calculate_power_flow from host buffers (host topology and plans, upload, run,
    # the output set of the step back to the host) on a leading sample of the scenarios
    e2e = None
    if args.e2e_fraction > 0:
        n_e2e = max(1, int(case.n_scenarios * args.e2e_fraction))
        upd = _leading(case.update, n_e2e)
        comps = selection if args.output == "timing" else \
            [c for c in ("node", "line", "sym_load", "asym_load") if c in case.input]
        t0 = time.perf_counter()
        model.calculate_power_flow(symmetric=case.symmetric, update_data=upd, factor_precision=args.precision,
                                   output_component_types=comps)
        secs = time.perf_counter() - t0
        e2e = {"value": n_e2e / secs, "unit": "scenarios/s", "scenarios": n_e2e, "seconds": secs,
               "outputs": comps}
        log(f"[bench] calculate_power_flow from host buffers: {n_e2e} scenarios in {secs:.2f}s")

    # adjoint gradients on the uneven-feeder batch (64 feeders, 128 scenarios): the solve, then one
    # adjoint pass (re-factorisation in FP64 at the converged state and the transposed sweeps)
    adjoint = None
    if not args.no_adjoint:
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
        del fb

    # automatic tap changer on year_20k_tap (the 110/10.5 kV transformer regulated): calculate_power_flow
    # from host buffers with a tap changing strategy; every round is one device batch over the scenarios
    # still searching. The oracle's tap changer runs beside it on a bounded sample.
    tap = None
    if not args.no_tap:
        tcase = synthetic.year_20k_tap()
        tmodel = PowerGridModel(tcase.input)
        tsel = dict(selection, transformer_tap_regulator=["tap_pos"])
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
            threads = len(os.sched_getaffinity(0))
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

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        rng = np.random.default_rng(0)
        # bounded sample of the same workload, about cpu_seconds of CPU work
        probe = rng.choice(case.n_scenarios, size=min(32, case.n_scenarios), replace=False)
        rate, _ = cpu_baseline(case, probe, threads)
        n_sample = int(min(case.n_scenarios, max(32, rate * args.cpu_seconds)))
        sample = rng.choice(case.n_scenarios, size=n_sample, replace=False)
        rate, secs = cpu_baseline(case, sample, threads)
        cpu = {"value": rate, "unit": "scenarios/s", "cores": threads, "kind": "port",
               "sample": f"{n_sample} random scenarios of {case.name} (FP64 CPU restatement, {threads} threads, "
                         f"{secs:.1f} s)"}
        log(f"[bench] cpu baseline {rate:.1f} scen/s on {threads} threads ({n_sample} scenarios, {secs:.1f}s)")

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
                   "output": ("columnar " + ", ".join(f"{c}{{{','.join(a)}}}" for c, a in selection.items()))
                   if args.output == "timing" else "whole records of every component"},
        "records_value": records_value,
        "fp64_gpu_value": fp64_value,
        "mixed_over_fp64": (value / fp64_value) if fp64_value else None,
        "roofline": {"bound": "hbm", "achieved": fwd_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fwd_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(case.name, args.precision),
                     "kernel": "forward_kernel (assembly + block LU + forward substitution)",
                     "bytes_per_launch": fwd_bytes_per_launch, "avg_launch_ms": st["fwd_launch_ms"]},
        "step_roofline": {"algorithmic_bytes": step_bytes, "achieved_gbs": step_bytes / t_step / 1e9,
                          "frac": step_bytes / t_step / 1e9 / HBM_PEAK_GBS},
        "pcie_inclusive": e2e,
        "adjoint": adjoint,
        "tap_changer": tap,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
