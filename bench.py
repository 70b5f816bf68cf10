#!/usr/bin/env python3
"""Throughput of the CNAB immersed-boundary step on one MI355X, in cell-steps/s.

Workload (default): a fixed cylinder (d=1, 628 points) and a flat plate (chord 1, 200 points)
pitching 30° ± 10° at ω = 1 on a 1024² finest grid with 5 levels, Re=200, dt=0.001, FP64.
A "step" is one advance! of that case; cell-steps/s = nx·ny·L·steps / seconds.  Inputs live in
HBM before the timed region.  Prints one JSON line (see the fields below).

The dominant kernel is the sine-transform contraction (MFMA).  The timed region replays the
captured step graphs.  Right after it, a few more steps of the same run are executed eagerly
with HIP events recorded around every contraction launch on the library's stream, and
`roofline` is computed from those events: algorithmic flops per launch divided by the average
launch time.  `roofline.traffic` is the per-launch HBM byte count measured for this workload
by rocprofv3 counter runs (profiles/contraction_traffic.json), or null when none was measured.
The CPU baseline runs the FP64 oracle (a C restatement of the reference, OpenMP) on a bounded
sample of the same workload.

The default run also measures the metric's other side on the batched sweep (`ellipse_batch`:
64 independent cases of 128²–512² cells advanced together): the mixed-precision batched step
(FP32 transforms, FP64 forces) and the FP64 batched step, with the mixed run's lift/drag error
against FP64, under config.mixed_precision_batched (skip with --no-batch).  `--workload
ellipse_batch` or `cylinder_array` (2048², 16 cylinders) bench those configurations alone.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "immersedboundary.jl_amd"))

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15 / 3, "mixed": 78.6e12}
# the contraction launches a profile reads: the precision's own (FP64 for the mixed form)
KIND = {"fp64": 0, "fp32": 1, "bf16x3": 2, "mixed": 0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, precision):
    """HBM bytes per contraction launch measured by rocprofv3 for this workload, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "contraction_traffic.json")) as f:
            return json.load(f).get(f"{workload}/{precision}", {}).get("bytes_per_launch")
    except OSError:
        return None


def by_shape(records, steps, peak, top=12):
    """The profiled contraction launches grouped by shape (equal algorithmic flops and bytes):
    launches per step, flops and bytes per launch, average device time, TFLOP/s and GB/s.
    Summing flops·launches over the groups and dividing by the summed time gives `achieved`."""
    us, fl, by = records
    groups = {}
    for u, f, b in zip(us, fl, by):
        g = groups.setdefault((f, b), [0, 0.0])
        g[0] += 1
        g[1] += u
    rows = []
    for (f, b), (n, t) in sorted(groups.items(), key=lambda kv: -kv[1][1]):
        avg = t / n
        rows.append({"launches_per_step": n / steps, "mflop_per_launch": f / 1e6,
                     "mbyte_per_launch": b / 1e6, "avg_us": avg,
                     "tflops": f / (avg * 1e-6) / 1e12, "frac_of_peak": f / (avg * 1e-6) / peak,
                     "gbps": b / (avg * 1e-6) / 1e9, "share_of_time": t / max(sum(us), 1e-30)})
    return rows[:top]


def cpu_baseline(make_problem, n_ref_iters, cap=8, threads=None):
    """Time the oracle on one step of the workload with cg! capped at `cap` iterations and on
    two lone B-products; extrapolate the step to `n_ref_iters` iterations of the reference's
    unpreconditioned cg! (the reference's force solve for moving bodies: it has no
    preconditioner).  `threads`: OpenMP threads of the oracle (default OMP_NUM_THREADS)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ibref

    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ibref.lib().ibr_set_threads(threads)
    prob = make_problem(force_solver="cg", cg_maxiter=cap)
    o = ibref.OracleRun(prob)
    t0 = time.perf_counter()
    o.advance(0.0)
    t_step = time.perf_counter() - t0
    z = o.Fb()
    t0 = time.perf_counter()
    for _ in range(2):
        o.B_times(z)
    t_B = (time.perf_counter() - t0) / 2
    t_fluid = max(t_step - (cap + 1) * t_B, 0.0)
    t_est = t_fluid + (n_ref_iters + 1) * t_B
    fl = prob.fluid
    cells = fl.nx * fl.ny * fl.nlevels
    o.close()
    return {
        "value": cells / t_est,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one oracle step of the same workload with cg! capped at {cap} iterations "
                   f"({t_step:.2f} s) and 2 lone B-products ({t_B:.3f} s each); step time "
                   f"extrapolated to the reference's unpreconditioned cg! count of "
                   f"{n_ref_iters} iterations on that step: {t_est:.1f} s/step"),
    }


def lift_drag_error(ibhip, make, nsteps, dt):
    """Run the workload in FP64 and in the requested precision side by side (untimed) and
    return the drag and lift errors: max over steps 2..n and bodies of |ΔCD| / max|CD| and
    |ΔCL| / max|CD| (CL is normalised by the drag scale because it is roundoff-seeded in
    symmetric cases; step 1 carries the impulsive-start spike and is reported apart)."""
    import numpy as np

    runs, fields = [], []
    for prec in ("fp64", None):
        prob = make(precision="fp64") if prec else make()
        st = ibhip.initstate(prob)
        sv = ibhip.initsolver(prob, st)
        cds, cls = [], []
        for k in range(nsteps):
            ibhip.advance_(st, sv, k * dt)
            cds.append(st.case.CD())
            cls.append(st.case.CL())
        runs.append((np.array(cds), np.array(cls)))
        c = st.case
        fields.append({"Γ": c.gamma(), "ψ": c.psi(), "q": c.q(), "nonlin1": c.nonlin(0),
                       "Fb": c.Fb()[:, None]})
        st.case.close()
    (cd0, cl0), (cd1, cl1) = runs
    scale = np.abs(cd0[1:]).max(axis=0)
    # every field at the last step, max-norm relative per level (level 1 first)
    ferr = {k: [float(np.abs(fields[1][k][:, j] - fields[0][k][:, j]).max() /
                      max(np.abs(fields[0][k][:, j]).max(), 1e-300))
                for j in range(fields[0][k].shape[1])] for k in fields[0]}
    return {"steps": nsteps,
            "CD_rel": float((np.abs(cd1[1:] - cd0[1:]).max(axis=0) / scale).max()),
            "CL_rel_to_CD": float((np.abs(cl1[1:] - cl0[1:]).max(axis=0) / scale).max()),
            "step1_CD_rel": float((np.abs(cd1[0] - cd0[0]) / np.abs(cd0[0])).max()),
            "fields_rel_per_level_at_last_step": ferr}


def batch_cpu_baseline(probs, every=8):
    """Time the oracle on a sample of the batch's cases (every `every`-th): one step with cg!
    capped at one iteration, minus its two B-products (the dense-inverse path applies a GEMV
    instead, whose cost is negligible here)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ibref

    from ibhip.problem import Problem, SchemeCNAB

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ibref.lib().ibr_set_threads(threads)
    cells = t_tot = 0.0
    picked = probs[::every]
    for p in picked:
        sch = p.scheme
        q = Problem(p.fluid, p.bodies, SchemeCNAB(sch.dt, precision="fp64", force_solver="cg",
                                                  cg_maxiter=1))
        o = ibref.OracleRun(q)
        o.advance(0.0)
        t0 = time.perf_counter()
        o.advance(sch.dt)  # a warm step
        t_step = time.perf_counter() - t0
        z = o.Fb()
        tb = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.B_times(z)
            tb.append(time.perf_counter() - t0)
        t_B = sorted(tb)[1]
        o.close()
        t_tot += max(t_step - 2 * t_B, 0.0)
        cells += p.fluid.nx * p.fluid.ny * p.fluid.nlevels
    return {
        "value": cells / t_tot,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the second oracle step of each of {len(picked)} cases (every {every}th of the "
                   f"batch, {int(cells)} cells in all) with cg! capped at one iteration, its two "
                   f"B-products timed apart and subtracted (the dense inverse of the batch path "
                   f"costs a GEMV): {t_tot:.2f} s for the sample"),
    }


def history_error(ref, run):
    """Drag and lift errors of one CD/CL history against another (arrays steps × cases), as
    in lift_drag_error: max over cases and steps 2..n."""
    import numpy as np

    (cd0, cl0), (cd1, cl1) = ref, run
    scale = np.abs(cd0[1:]).max(axis=0)
    return {"steps": int(cd0.shape[0]),
            "CD_rel": float((np.abs(cd1[1:] - cd0[1:]).max(axis=0) / scale).max()),
            "CL_rel_to_CD": float((np.abs(cl1[1:] - cl0[1:]).max(axis=0) / scale).max()),
            "step1_CD_rel": float((np.abs(cd1[0] - cd0[0]) / np.abs(cd0[0])).max())}


def measure_batch(torch, ibhip, precision, steps, warmup, profile_steps, acc_steps=20):
    """The batched form: the ellipse sweep (64 independent cases of 128²–512² cells, 2–5
    levels) advanced together, one grouped launch per substep, the step replayed from a HIP
    graph.  A step is one advance! of every case; cell-steps/s counts all cases.  The first
    acc_steps steps (untimed, at least the warmup) record every case's CD and CL."""
    import numpy as np

    from ibhip.workloads import ellipse_batch

    make = lambda **kw: ellipse_batch(**{"precision": precision, **kw})  # noqa: E731
    probs = make()
    cells = sum(p.fluid.nx * p.fluid.ny * p.fluid.nlevels for p in probs)
    npts = sum(p.bodies.npoints for p in probs)
    log(f"bench: ellipse_batch {len(probs)} cases, {cells} cells/step, {npts} points, "
        f"{precision}")
    t0 = time.perf_counter()
    states = [ibhip.initstate(p) for p in probs]
    for p, st in zip(probs, states):
        ibhip.initsolver(p, st)
    batch = ibhip.Batch(states)
    batch.sync()
    t_setup = time.perf_counter() - t0
    log(f"bench: setup {t_setup:.2f} s")
    warm = max(warmup, 3, acc_steps)  # step 1 tunes, steps 2 and 3 capture the two graphs
    cds, cls = [], []
    for _ in range(warm):
        batch.advance()
        batch.sync()
        cds.append(np.array([st.case.CD()[0] for st in states]))
        cls.append(np.array([st.case.CL()[0] for st in states]))
    marker = torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        batch.advance()
        if (s + 1) % 25 == 0:
            log(f"bench: {s + 1}/{steps} steps")
    batch.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    marker.zero_()
    torch.cuda.synchronize()
    batch.profile(True)
    t1 = time.perf_counter()
    for _ in range(profile_steps):
        batch.advance()
    batch.sync()
    prof_elapsed = time.perf_counter() - t1
    kind = KIND[precision]
    ms, launches, flops = batch.profile_read(kind)
    ev_ms, ev_launches, _ = batch.profile_read(kind, events=True)
    records = batch.profile_launches(kind)
    batch.profile(False)
    launches_per_step = batch.launches_per_step()
    tune = batch.tune_report()
    batch.close()
    for st in states:
        st.case.close()
    avg_launch_s = (ms / 1e3) / max(launches, 1)
    achieved = (flops / max(launches, 1)) / avg_launch_s if launches else 0.0
    return {
        "probs": probs, "cells": cells, "npts": npts, "warmup": warm, "elapsed": elapsed,
        "value": cells * steps / elapsed, "ms_per_step": 1e3 * elapsed / max(steps, 1),
        "setup_s": t_setup, "launches_per_step": launches_per_step, "tune": tune,
        "history": (np.array(cds), np.array(cls)),
        "roofline": {
            "bound": "mfma",
            "achieved": achieved / 1e12,
            "peak": PEAK[precision] / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK[precision],
            "traffic": measured_traffic("ellipse_batch", precision),
            "kernel": f"kg_gemm_eo<{precision}> (grouped sine-transform contraction)",
            "avg_launch_us": avg_launch_s * 1e6,
            "event_avg_launch_us": 1e3 * ev_ms / max(ev_launches, 1),
            "launches": launches,
            "gemm_share_of_profiled_steps": (ms / 1e3) / prof_elapsed,
            "by_shape": by_shape(records, profile_steps, PEAK[precision]),
            "timing": "device wall clock of each grouped contraction launch (first workgroup "
                      "start to last workgroup end) over a profiled pass of eager steps right "
                      "after the timed (graph) region; event_avg_launch_us: HIP events around "
                      "each launch on the batch stream",
        },
    }


DTYPE = {"fp64": "f64", "fp32": "f32 transforms / f64 forces",
         "bf16x3": "bf16x3 transforms / f64 forces",
         "mixed": "f64, f32 transforms on the final Vort2Flux pass's coarse levels"}


def run_batch(args, torch, ibhip):
    m = measure_batch(torch, ibhip, args.precision, args.steps, args.warmup, args.profile_steps)
    m["accuracy"] = None
    if args.precision != "fp64":
        ref = measure_batch(torch, ibhip, "fp64", 0, 0, 0)
        m["accuracy"] = history_error(ref["history"], m["history"])
        log(f"bench: lift/drag error vs FP64 {m['accuracy']}")
    out = {
        "metric": "cell-steps/s: FP64 vs mixed-precision batched, at ≤1e-6 rel lift/drag err",
        "value": m["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": m["warmup"],
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": (f"ellipse_batch: {len(m['probs'])} independent cases, grids "
                                f"128²–512², 2–5 levels, one fixed ellipse each ({m['npts']} "
                                f"points in all), Re 50–400, dt = 0.2h; {m['cells']} cells per "
                                f"step"),
                   "cases": len(m["probs"]), "cells_per_step": m["cells"],
                   "body_points": m["npts"], "precision": args.precision,
                   "setup_s": m["setup_s"], "grouped_launches_per_step": m["launches_per_step"],
                   "lift_drag_error_vs_fp64": m["accuracy"],
                   "autotuned_contractions": m["tune"]},
        "roofline": m["roofline"],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = batch_cpu_baseline(m["probs"])
        log("bench: cpu baseline done")
    print(json.dumps(out))


def batched_summary(torch, ibhip, args, plain_value, plain_steady=None):
    """The metric's comparison on the 64-case sweep, beside the plain FP64 single-case step
    measured above (its timed window, and the continuation with its preconditioner rebuilds):
    the FP64 batched step, the mixed batched step (FP64 where the transform
    error reaches the forces, FP32 on the final pass's coarse levels) and the FP32 batched
    step (FP32 transforms refined against FP64 residuals); CD/CL errors against the FP64
    batch over the first 20 steps."""
    m = {p: measure_batch(torch, ibhip, p, args.batch_steps, 3, 2)
         for p in ("fp64", "mixed", "fp32")}
    out = {
        "workload": (f"ellipse_batch: {len(m['fp64']['probs'])} cases, 128²–512², 2–5 levels, "
                     f"{m['fp64']['cells']} cells per step"),
        "steps": args.batch_steps,
    }
    for p in ("fp64", "mixed", "fp32"):
        out[f"{p}_cell_steps_per_s"] = m[p]["value"]
        out[f"{p}_ms_per_step"] = m[p]["ms_per_step"]
        out[f"{p}_batched_vs_plain_fp64_step"] = m[p]["value"] / plain_value
        if plain_steady:  # against the plain step's rate with its preconditioner rebuilds
            out[f"{p}_batched_vs_plain_fp64_steady"] = m[p]["value"] / plain_steady
        if p != "fp64":
            out[f"{p}_vs_fp64_batched"] = m[p]["value"] / m["fp64"]["value"]
            out[f"{p}_lift_drag_error_vs_fp64"] = history_error(m["fp64"]["history"],
                                                                 m[p]["history"])
        out[f"{p}_contraction_roofline"] = {k: m[p]["roofline"][k] for k in
                                            ("achieved", "peak", "frac", "avg_launch_us")}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps (the moving-body preconditioner is rebuilt inside the "
                         "window as often as the run needs, so its cost is amortized)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cylinder_plate")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3", "mixed"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="steps of the profiled (eager) pass that times the contraction kernel")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched (64-case) FP32/FP64 comparison of the default run")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--steady-steps", type=int, default=300,
                    help="moving bodies: steps of the same run after the timed window, timed "
                         "with the preconditioner rebuilds they need (0: skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one elementwise copy of 2^25 doubles first (256 MiB read + 256 MiB "
                         "written) so that profiler byte counters can be scaled")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); independent cases shard across GPUs with no communication",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import ibhip
    from ibhip.workloads import WORKLOADS

    if args.calib_copy:
        x = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y
    if args.workload == "ellipse_batch":
        run_batch(args, torch, ibhip)
        return
    make = lambda **kw: WORKLOADS[args.workload](**{"precision": args.precision, **kw})  # noqa: E731
    prob = make()
    fl = prob.fluid
    cells = fl.nx * fl.ny * fl.nlevels
    dt = prob.scheme.dt
    log(f"bench: {args.workload} {fl.nx}x{fl.ny} L={fl.nlevels} nb={prob.bodies.npoints} "
        f"{args.precision}")

    t0 = time.perf_counter()
    state = ibhip.initstate(prob)
    solver = ibhip.initsolver(prob, state)
    case = state.case
    case.sync()
    t_setup = time.perf_counter() - t0
    log(f"bench: setup {t_setup:.2f} s")
    k = 0
    for _ in range(args.warmup):
        ibhip.advance_(state, solver, k * dt)
        k += 1
    case.sync()
    log(f"bench: warmup done ({args.warmup} steps)")

    iters = []
    st0 = case.solver_stats()
    marker = torch.zeros(1, device="cuda")  # a fill kernel that marks the timed window in traces
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ibhip.advance_(state, solver, k * dt)
        iters.append(case.cg_iters())
        k += 1
        if (s + 1) % 10 == 0:
            log(f"bench: {s + 1}/{args.steps} steps")
    case.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st1 = case.solver_stats()
    marker.zero_()
    torch.cuda.synchronize()
    # per-launch timing of the contraction kernels: the library records HIP events around
    # each launch on its stream; with events on, steps run as eager launches (no graph), so
    # this is a separate pass over the next steps of the same run
    case.profile(True)
    t1 = time.perf_counter()
    for s in range(args.profile_steps):
        ibhip.advance_(state, solver, k * dt)
        k += 1
    case.sync()
    prof_elapsed = time.perf_counter() - t1
    kind = KIND[args.precision]
    ms, launches, flops = case.profile_read(kind)
    ev_ms, ev_launches, _ = case.profile_read(kind, events=True)
    records = case.profile_launches(kind)
    case.profile(False)
    log(f"bench: profiled pass {args.profile_steps} steps")
    steady = None
    if prob.any_moving and args.steady_steps > 0:
        # the same run continued for steady_steps more steps: the preconditioner is rebuilt
        # as often as the bodies' motion asks, which a short timed window may not show
        s0 = case.solver_stats()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(args.steady_steps):
            ibhip.advance_(state, solver, k * dt)
            k += 1
            if (s + 1) % 100 == 0:
                log(f"bench: steady state {s + 1}/{args.steady_steps} steps")
        case.sync()
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        s1 = case.solver_stats()
        steady = {"steps": args.steady_steps, "ms_per_step": 1e3 * el / args.steady_steps,
                  "cell_steps_per_s": cells * args.steady_steps / el,
                  "pc_rebuilds": s1["pc_rebuilds"] - s0["pc_rebuilds"],
                  "pc_rebuild_ms_per_step": 1e3 * (s1["pc_rebuild_s"] - s0["pc_rebuild_s"])
                  / args.steady_steps}
        log(f"bench: steady state {steady}")
    accuracy = None
    if args.precision != "fp64":
        accuracy = lift_drag_error(ibhip, make, args.warmup + args.steps, dt)
        log(f"bench: lift/drag error vs FP64 {accuracy}")
    ms_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed

    avg_launch_s = (ms / 1e3) / max(launches, 1)
    achieved = (flops / max(launches, 1)) / avg_launch_s if launches else 0.0
    peak = PEAK["fp64"] if args.precision == "fp64" else PEAK[args.precision]
    traffic = measured_traffic(args.workload, args.precision)
    roofline = {
        "bound": "mfma",
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": traffic,
        "kernel": f"k_gemm_eo<{args.precision}> (split sine-transform contraction)",
        "avg_launch_us": avg_launch_s * 1e6,
        "launches": launches,
        "event_avg_launch_us": 1e3 * ev_ms / max(ev_launches, 1),
        "event_launches": ev_launches,
        "gemm_share_of_profiled_steps": (ms / 1e3) / prof_elapsed,
        "by_shape": by_shape(records, args.profile_steps, peak),
        "timing": "device wall clock of each contraction launch (first workgroup start to last "
                  "workgroup end, as rocprofv3's kernel trace measures it) over a profiled pass "
                  "of eager steps right after the timed (graph) region; launches skipped on the "
                  "device after the CG converged are left out of time, count and flops. "
                  "event_avg_launch_us: HIP events around every launch on the library stream "
                  "(adds the dispatch gap, counts the skipped launches)",
    }
    cg = {"force_solver": "pcg" if prob.any_moving else "dense",
          "iters_mean": sum(iters) / max(len(iters), 1), "iters_max": max(iters or [0])}
    if prob.any_moving:
        # preconditioner rebuilds inside the timed steps (moving bodies' columns + inverse)
        cg.update({"pc_rebuilds_timed": st1["pc_rebuilds"] - st0["pc_rebuilds"],
                   "pc_rebuild_ms_per_step": 1e3 * (st1["pc_rebuild_s"] - st0["pc_rebuild_s"])
                   / args.steps,
                   "pc_columns_ms_total": 1e3 * (st1["pc_columns_s"] - st0["pc_columns_s"]),
                   "pc_inverse_ms_total": 1e3 * (st1["pc_inverse_s"] - st0["pc_inverse_s"]),
                   "iteration_ms": 1e3 * st1["iteration_s"],
                   "low_rank_coarse_part": case.lowrank_stats(),
                   "steady_state": steady})
    out = {
        "metric": "cell-steps/s: FP64 vs mixed-precision batched, at ≤1e-6 rel lift/drag err",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {fl.nx}x{fl.ny} cells, {fl.nlevels} levels, "
                               f"{prob.bodies.npoints} body points, Re={fl.Re:g}, dt={dt:g}",
                   "nx": fl.nx, "ny": fl.ny, "levels": fl.nlevels,
                   "body_points": prob.bodies.npoints, "precision": args.precision,
                   "finest_only_cell_steps_per_s": value / fl.nlevels,
                   "setup_s": t_setup, "force_solve": cg,
                   "lift_drag_error_vs_fp64": accuracy,
                   "autotuned_contractions": case.tune_report()},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        # iteration count of the reference algorithm (unpreconditioned cg!) on the first step
        ref_prob = make(force_solver="cg")
        rs = ibhip.initstate(ref_prob)
        rv = ibhip.initsolver(ref_prob, rs)
        ibhip.advance_(rs, rv, 0.0)
        n_ref = rs.case.cg_iters()
        rs.case.close()
        log(f"bench: reference cg! iterations on step 1: {n_ref}")
        out["cpu_baseline"] = cpu_baseline(make, n_ref)
        one = cpu_baseline(make, n_ref, threads=1)
        out["cpu_baseline"]["one_core"] = {"value": one["value"], "sample": one["sample"]}
        log("bench: cpu baseline done")
    if not args.no_batch and args.workload == "cylinder_plate" and args.precision == "fp64":
        out["config"]["mixed_precision_batched"] = batched_summary(
            torch, ibhip, args, value, steady["cell_steps_per_s"] if steady else None)
        log("bench: batched comparison done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
