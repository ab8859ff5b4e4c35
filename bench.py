#!/usr/bin/env python3
"""Benchmark of the hydraulic-head path on one MI355X.

Workload: a synthetic glacier on a 4096 x 1024 grid (100 km x 20 km; bed
Gaussian random field of 50 m standard deviation, 2 km correlation length;
SHMIP-like surface; OCEAN column at x = 0), harmonic K_face, implicit gap
scheme, regularised-Coulomb sliding (C = 0.5), seasonal melt (1e-8 m/s
baseline plus a 1e-7 m/s one-day sinusoid) plus 20 seeded moulins of
1-10 m^3/s, dt = 3600 s; each timestep runs the Picard loop to the 1e-6
relative tolerance with Chebyshev(4)-preconditioned CG (every solve to a
relative preconditioned residual of 1e-11), then the gap step.

With ``--workload parabolic`` the run is the parabolic head scheme instead:
8192 x 2048 cells, bed amplitude 100 m plus a seeded 2 km-wide trough,
e_v = 1e-3, dt = 600 s, one backward-Euler solve per timestep.

``--forms`` picks the solver forms run from the same initial state (plain
FP64 Chebyshev(4)-PCG; FP32 inner solves with FP64 refinement; FP64 CG with
the FP32 multigrid V-cycle; the same warm-started by the U-Net); ``value`` is
the fastest form's time and every other form is reported beside it with its
per-field error against the FP64 form.

A step is one whole timestep.  W untimed warm-up steps (the first solves the
initial-condition transient), then K timed steps bracketed by device
synchronisation.  The JSON line reports seconds per timestep, the roofline of
the dominant kernel (algorithmic bytes per launch / mean launch time from HIP
events on the library's stream), the operator apply's bandwidth, and the CPU
baseline: the FP64 restatement of the same algorithm timed on the host's
cores on a bounded sample and scaled by the GPU run's per-step CG and Picard
counts.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "shakti.jl_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(nx, ny, bed_amp, seed, trough=0.0, trough_depth=0.0):
    from shakti_hip import synthetic as S
    g, raw = S.glacier(nx, ny, bed_amp=bed_amp, seed=seed, trough_width=trough, trough_depth=trough_depth)
    moulin = S.moulin_field(g, raw["mask"], n=20, seed=seed)
    return g, raw, moulin


class _ParabolicCounts:
    """PicardSolver-like view of a parabolic run: one linear solve per step."""

    def __init__(self, ls):
        self.ls = ls
        self.converged = True
        self.last_iter = 1

    @property
    def cg_iters(self):
        return self.ls.ws.stats.niter


def gpu_run(args, g, raw, moulin, form):
    import shakti_hip as H
    from shakti_hip import synthetic as S
    parabolic = args.workload == "parabolic"
    p = H.ModelParameters(e_v=args.e_v) if parabolic else H.ModelParameters()
    sl = H.RegularizedCoulombSlidingLaw(0.5)
    s = H.State(g, p)
    H.set_initial_conditions(s, g, p, sl, **raw)
    # fp64: the reference's Chebyshev(4)-PCG in FP64; fp32_refine: FP32 Chebyshev(4)-PCG inside FP64
    # refinement; fp32_mg: FP32 multigrid-PCG inside FP64 refinement; fp64_mg: FP64 CG preconditioned
    # by the FP32 multigrid V-cycle; learned_mg: fp64_mg from the U-Net's warm start
    if form == "fp64":
        kw = dict(chebyshev_degree=4, rtol=args.rtol, form="fp64")
    elif form == "fp64_mg":
        kw = dict(multigrid=True, rtol=args.mg_rtol, form="fp64")
    elif form == "learned_mg":
        # fp64_mg warm-started by the trained U-Net (its own forward kernels, BF16 on MFMA)
        from shakti_hip import learned as LN
        H8, W8 = LN.canvas_shape(g.nx, g.ny)
        net = LN.UNet.load(args.weights, 1, H8, W8)
        kw = dict(multigrid=True, rtol=args.mg_rtol, form="learned", network=net)
    elif form == "fp32_refine":
        kw = dict(chebyshev_degree=4, rtol=args.ir_rtol, form="fp32_refine", inner_rtol=args.inner_rtol)
    else:
        kw = dict(multigrid=True, rtol=args.ir_rtol, form="fp32_refine", inner_rtol=args.mg_inner_rtol)
    ls = H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, atol=0.0, check_every=args.check_every, **kw)
    mi = H.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
    if parabolic:
        # ParabolicHeadScheme: one backward-Euler solve per step (src/parabolic_solver.jl:16-38)
        ps = _ParabolicCounts(ls)
        sim = H.Simulation(g, s, args.warmup + args.steps, args.dt, p, "implicit", [], mi, sl, ls=ls,
                           k_face_choice="harmonic")
    else:
        ps = H.PicardSolver(args.max_picard, 1e-6, ls, g)
        sim = H.Simulation(g, s, args.warmup + args.steps, args.dt, p, "implicit", [], mi, sl, ps=ps,
                           k_face_choice="harmonic")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        H.step(sim)
        sim.total_time += sim.dt
        log(f"[bench {form}] warm-up step {w + 1}/{args.warmup}: {time.perf_counter() - t0:.3f} s, "
            f"Picard {ps.last_iter} (converged {ps.converged}), CG {ps.cg_iters}")
    times, picard, cg, conv = [], [], [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        H.step(sim)  # returns after the library's stream has synchronised
        dt = time.perf_counter() - t0
        sim.total_time += sim.dt
        times.append(dt)
        picard.append(ps.last_iter)
        cg.append(ps.cg_iters)
        conv.append(ps.converged)
        log(f"[bench {form}] step {k + 1}/{args.steps}: {dt:.3f} s, Picard {ps.last_iter} "
            f"(converged {ps.converged}), CG {ps.cg_iters}")
    # Per-kernel durations: one more step with HIP events bracketing every
    # launch on the library stream (CG batches are then launched one by one
    # instead of as a captured graph, so this step is not part of the timing).
    s.profile_reset()
    s.profile_enable(None)
    t0 = time.perf_counter()
    H.step(sim)
    prof_step = time.perf_counter() - t0
    sim.total_time += sim.dt
    prof = s.profile_read()
    s.profile_enable([])
    log(f"[bench {form}] profiled step: {prof_step:.3f} s, Picard {ps.last_iter}, CG {ps.cg_iters}")
    # operator apply alone, back-to-back launches on the library stream
    import ctypes as C
    from shakti_hip import _lib as L
    mean_ms, nbytes = C.c_double(), C.c_double()
    L.check(L.load().shk_time_kernel(s.ctx, 0, 50, C.byref(mean_ms), C.byref(nbytes)), s.ctx, "time_kernel")
    op = {"kernel": "operator apply (FP64, symmetric face layout)", "mean_ms": mean_ms.value,
          "bytes_per_launch": nbytes.value, "achieved_GBs": nbytes.value / (mean_ms.value * 1e-3) / 1e9}
    op["frac"] = op["achieved_GBs"] / HBM_PEAK_GBS
    L.check(L.load().shk_time_kernel(s.ctx, 2, 50, C.byref(mean_ms), C.byref(nbytes)), s.ctx, "time_kernel")
    op["copy_reference_GBs"] = nbytes.value / (mean_ms.value * 1e-3) / 1e9
    # cold: each launch alone after a 1 GiB read has flushed L2 and the Infinity Cache
    L.check(L.load().shk_time_kernel(s.ctx, 3, 10, C.byref(mean_ms), C.byref(nbytes)), s.ctx, "time_kernel")
    op["cold_mean_ms"] = mean_ms.value
    op["cold_GBs"] = nbytes.value / (mean_ms.value * 1e-3) / 1e9
    op["cold_frac"] = op["cold_GBs"] / HBM_PEAK_GBS
    L.check(L.load().shk_time_kernel(s.ctx, 4, 10, C.byref(mean_ms), C.byref(nbytes)), s.ctx, "time_kernel")
    op["cold_copy_reference_GBs"] = nbytes.value / (mean_ms.value * 1e-3) / 1e9
    # the FP64 CG matvec kernel as the solve runs it: each launch timed alone
    # after a 1 GiB read, since in the solve the V-cycle between two matvecs
    # streams their inputs out of L2 and the Infinity Cache
    L.check(L.load().shk_time_kernel(s.ctx, 5, 20, C.byref(mean_ms), C.byref(nbytes)), s.ctx, "time_kernel")
    op["cg_matvec_mean_ms"] = mean_ms.value
    op["cg_matvec_bytes_per_launch"] = nbytes.value
    fields = {f: s.get(f) for f in ("h", "b", "N", "q_x", "q_y", "mdot", "K")}
    return dict(times=times, picard=picard, cg=cg, conv=conv, prof=prof, prof_step=prof_step,
                prof_cg=ps.cg_iters, operator=op, fields=fields, state=s)


def cpu_baseline(args, g, raw, moulin, picard_per_step, cg_per_step):
    """The oracle (FP64 C restatement, OpenMP) on the same inputs: one solve
    capped at 1 CG iteration, one at 1 + S, and one derived refresh, timed on
    the host; scaled by the GPU run's per-step Picard and CG counts."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.build()
    og = O.Grid(g.nx, g.ny, g.lx, g.ly)
    par = args.workload == "parabolic"
    p = O.ModelParameters(e_v=args.e_v) if par else O.ModelParameters()
    m = O.Model(kfs=1, sliding_C=0.5)
    s = O.State(og)
    O.set_initial_conditions(s, og, p, m, **raw)
    s.ieb[...] = s.ieb + moulin
    S = args.cpu_iters
    sv1 = O.Solver(og, "cg", precond="chebyshev", degree=4, atol=0.0, rtol=1e-30, itmax=1)
    h0 = s.h.copy()

    def solve(sv):
        if par:
            sv.solve_parabolic(s, p, 1, args.dt)
        else:
            sv.solve_elliptic(s, p, 1)

    t0 = time.perf_counter()
    solve(sv1)
    t1 = time.perf_counter() - t0
    s.h[...] = h0
    svS = O.Solver(og, "cg", precond="chebyshev", degree=4, atol=0.0, rtol=1e-30, itmax=S + 1)
    t0 = time.perf_counter()
    solve(svS)
    tS = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.update_derived(s, og, p, m)
    td = time.perf_counter() - t0
    per_cg = (tS - t1) / S
    per_picard = (t1 - per_cg) + td
    est = cg_per_step * per_cg + picard_per_step * per_picard
    sample = (f"oracle on {g.nx}x{g.ny}: 1 solve capped at 1 CG iteration ({t1:.2f} s), 1 capped at {S + 1} "
              f"({tS:.2f} s), 1 derived refresh ({td:.2f} s); {per_cg * 1e3:.1f} ms per Chebyshev(4)-PCG "
              f"iteration, {per_picard:.2f} s per {'solve' if par else 'Picard iteration'} outside CG; scaled "
              f"by the GPU run's {picard_per_step:.1f} {'solves' if par else 'Picard iterations'} and "
              f"{cg_per_step:.0f} CG iterations per step")
    return {"value": est, "unit": "s/timestep", "cores": O.num_threads(), "kind": "port", "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["elliptic", "parabolic"], default="elliptic",
                    help="elliptic: Picard to 1e-6 per step (4096x1024 defaults); parabolic: one backward-Euler "
                         "solve per step, e_v 1e-3, dt 600 s, 8192x2048, bed amplitude 100 m and a 2 km trough")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--bed-amp", type=float, default=None)
    ap.add_argument("--dt", type=float, default=None)
    ap.add_argument("--e-v", type=float, default=1e-3)
    ap.add_argument("--trough-depth", type=float, default=200.0,
                    help="parabolic workload: depth (m) of the seeded 2 km-wide trough")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rtol", type=float, default=1e-11)
    ap.add_argument("--check-every", type=int, default=32)
    ap.add_argument("--forms", default="fp64,fp32_mg,fp64_mg,learned_mg",
                    help="comma-separated solver forms to run from the same initial state")
    ap.add_argument("--ir-rtol", type=float, default=1e-8,
                    help="fp32_refine: FP64 residual reduction per linear solve")
    ap.add_argument("--inner-rtol", type=float, default=1e-3,
                    help="fp32_refine: residual reduction of each FP32 inner solve")
    ap.add_argument("--mg-inner-rtol", type=float, default=1e-4,
                    help="fp32_mg: residual reduction of each FP32 multigrid-PCG inner solve")
    ap.add_argument("--mg-rtol", type=float, default=1e-10,
                    help="fp64_mg: reduction of the multigrid-preconditioned residual norm per linear solve")
    ap.add_argument("--weights", default=str(ROOT / "shakti.jl_amd" / "shakti_hip" / "weights" /
                                             "unet_16_32_64_128.npz"),
                    help="learned_mg: trained U-Net parameters")
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--max-picard", type=int, default=500,
                    help="Picard iteration cap per step (500 as in the reference configs; lower only for profiling)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    par = args.workload == "parabolic"
    args.nx = args.nx or (8192 if par else 4096)
    args.ny = args.ny or (2048 if par else 1024)
    args.bed_amp = args.bed_amp if args.bed_amp is not None else (100.0 if par else 50.0)
    args.dt = args.dt or (600.0 if par else 3600.0)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the head solve is not sharded across GPUs "
              "(see the sharding design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    log(f"[bench] building inputs {args.nx}x{args.ny}, bed amplitude {args.bed_amp} m, seed {args.seed}")
    g, raw, moulin = make_inputs(args.nx, args.ny, args.bed_amp, args.seed, trough=2000.0 if par else 0.0,
                                 trough_depth=args.trough_depth if par else 0.0)
    forms = args.forms.split(",")
    runs = {}
    for form in forms:
        runs[form] = gpu_run(args, g, raw, moulin, form)
        runs[form]["state"] = None
    K = args.steps
    form_summary = {}
    for form, rr in runs.items():
        form_summary[form] = {"s_per_timestep": round(sum(rr["times"]) / K, 4),
                              "picard_iters_per_step": sum(rr["picard"]) / K,
                              "cg_iters_per_step": sum(rr["cg"]) / K,
                              "picard_converged_all": all(rr["conv"])}
    if "fp64" in runs:
        import numpy as np
        ref = runs["fp64"]["fields"]
        for form, rr in runs.items():
            if form == "fp64":
                continue
            errs = {}
            for f, a in rr["fields"].items():
                scale = float(np.max(np.abs(ref[f])))
                errs[f] = float(np.max(np.abs(a - ref[f]))) / scale if scale > 0 else float(np.max(np.abs(a)))
            form_summary[form]["max_rel_err_vs_fp64"] = errs
            form_summary[form]["speedup_vs_fp64"] = round(form_summary["fp64"]["s_per_timestep"] /
                                                          form_summary[form]["s_per_timestep"], 3)
    best = min(runs, key=lambda f: form_summary[f]["s_per_timestep"])
    r = runs[best]
    tot = sum(r["times"])
    s_per_step = tot / K
    picard_per_step = sum(r["picard"]) / K
    cg_per_step = sum(r["cg"]) / K
    prof = r["prof"]
    # dominant kernel: largest total time among the profiled CG kernels
    dom_name = max((k for k in prof if prof[k]["bytes"] > 0), key=lambda k: prof[k]["ms"])
    d = prof[dom_name]
    event_mean_ms = d["ms"] / d["count"]  # each launch bracketed alone by events (includes its dispatch)
    mean_ms = event_mean_ms
    bytes_per_launch = d["bytes"] / d["count"]
    how = "profiled step, each launch bracketed by HIP events on the library stream"
    fp64_cg = best in ("fp64", "fp64_mg", "learned_mg")  # the forms whose CG matvec is the FP64 kernel
    if dom_name == "cg_matvec" and fp64_cg and "cg_matvec_mean_ms" in r["operator"]:
        # each launch alone between HIP events on the library stream, after a
        # 1 GiB read (the cache state the solve gives it): comparable with the
        # profiler's mean over the solve's active launches
        mean_ms = r["operator"]["cg_matvec_mean_ms"]
        bytes_per_launch = r["operator"]["cg_matvec_bytes_per_launch"]
        how = "20 launches on the library stream, each alone between HIP events after a 1 GiB read flush"
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    traffic = None
    rocprof_ms = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            t = json.loads(tfile.read_text())
            key = f"{dom_name}@{args.nx}x{args.ny}"
            if key in t:
                traffic = t[key]["hbm_bytes_per_launch"]
                rocprof_ms = t[key].get("mean_us_rocprof", 0.0) * 1e-3 or None
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": bytes_per_launch, "mean_launch_ms": mean_ms, "timing": how,
                "event_mean_ms_profiled_step": event_mean_ms, "rocprof_mean_ms": rocprof_ms,
                "launches": d["count"]}
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle restatement) on a bounded sample ...")
        try:
            base = runs.get("fp64", r)  # the CPU restatement runs the plain FP64 algorithm
            cpu = cpu_baseline(args, g, raw, moulin, sum(base["picard"]) / K, sum(base["cg"]) / K)
        except Exception as e:  # the baseline is reported, never required
            log(f"[bench] CPU baseline failed: {e!r}")
    kernels = {k: {"launches": v["count"], "mean_ms": v["ms"] / v["count"],
                   "GBs": (v["bytes"] / v["count"]) / (v["ms"] / v["count"] * 1e-3) / 1e9 if v["bytes"] else None,
                   "share_of_profiled_step": v["ms"] * 1e-3 / r["prof_step"]} for k, v in prof.items()}
    out = {
        "metric": "s per timestep to 1e-6 Picard tol & HBM GB/s: mixed/learned vs plain FP64 PCG",
        "value": round(s_per_step, 4),
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_step * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"{args.nx}x{args.ny} synthetic glacier, bed amplitude {args.bed_amp} m, harmonic "
                                f"K_face, Chebyshev(4)-PCG rtol {args.rtol}, Picard tol 1e-6, implicit gap, "
                                "seasonal + 20 moulins, dt 3600 s") if not par else
                               (f"{args.nx}x{args.ny} synthetic glacier, bed amplitude {args.bed_amp} m + 2 km trough "
                                f"{args.trough_depth} m deep, parabolic scheme e_v {args.e_v}, one backward-Euler "
                                f"solve per step, harmonic K_face, implicit gap, seasonal + 20 moulins, "
                                f"dt {args.dt} s"),
                   "form": best, "nx": args.nx, "ny": args.ny, "cells": args.nx * args.ny},
        "forms": form_summary,
        "timesteps_per_s": round(1.0 / s_per_step, 4),
        "picard_iters_per_step": picard_per_step,
        "picard_iters_per_s": round(picard_per_step / s_per_step, 3),
        "cg_iters_per_step": cg_per_step,
        "picard_converged_all": all(all(rr["conv"]) for rr in runs.values()),
        "roofline": roofline,
        "profiled_step": {"seconds": r["prof_step"], "cg_iters": r["prof_cg"],
                          "note": "one extra step after the timed ones, every launch bracketed by HIP events"},
        "operator": r["operator"],
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
