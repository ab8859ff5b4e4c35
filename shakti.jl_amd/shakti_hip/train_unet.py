"""Train the U-Net warm start on pairs produced by the solver's own runs.

    python -m shakti_hip.train_unet [--glaciers 64] [--per-glacier 320] [--iters 3000] [--out PATH]

1. Runs the plain solver (FP64 CG, multigrid preconditioner) over seeded
   512x128 synthetic glaciers (bed amplitude and correlation length varied)
   and records one (network input, scaled Picard correction) pair per Picard
   iteration, on the device.
2. Trains the network with the library's own forward / backward / Adam
   kernels on random batches (the last glaciers are held out for validation).
3. Saves the FP32 parameters (npz) and prints one JSON line with the losses,
   the timings and what the trained warm start gains on held-out glaciers
   (CG iterations and seconds per timestep with and without it).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api as H
from . import learned as LN
from . import synthetic as S


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_steps(g, raw, moulin, form: str, net, steps: int, precond: str = "multigrid", rtol: float = 1e-10,
              every: int = 0):
    """Timesteps of the 1024x256-style workload on glacier (g, raw); returns
    per-step (seconds, Picard iterations, CG iterations) and the final h.
    every > 0 (form "learned"): the network also preconditions one CG
    iteration in `every`."""
    p = H.ModelParameters()
    sl = H.RegularizedCoulombSlidingLaw(0.5)
    s = H.State(g, p)
    H.set_initial_conditions(s, g, p, sl, **raw)
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    ls = H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, atol=0.0, rtol=rtol, check_every=16, form=form,
                             network=net if form == "learned" else None,
                             learned_precond_every=every if form == "learned" else 0, **kw)
    ps = H.PicardSolver(500, 1e-6, ls, g)
    mi = H.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
    sim = H.Simulation(g, s, steps, 3600.0, p, "implicit", [], mi, sl, ps=ps, k_face_choice="harmonic")
    out = []
    for _ in range(steps):
        t0 = time.perf_counter()
        H.step(sim)
        out.append((time.perf_counter() - t0, ps.last_iter, ps.cg_iters))
        sim.total_time += sim.dt
    return out, s.get("h")


def gain(net_path, nx: int, ny: int, seeds, steps: int, precond: str = "multigrid", every: int = 0):
    """CG iterations and seconds per timestep with and without the learned
    warm start (and, every > 0, the learned preconditioner), on glaciers the
    network never saw."""
    H8, W8 = LN.canvas_shape(nx, ny)
    net = LN.UNet.load(net_path, 1, H8, W8)
    rows = []
    for sd in seeds:
        g, raw = S.glacier(nx, ny, bed_amp=50.0, seed=sd)
        moulin = S.moulin_field(g, raw["mask"], n=20, seed=sd)
        res = {}
        for form in ("fp64", "learned"):
            res[form] = run_steps(g, raw, moulin, form, net, steps, precond, every=every)
        base, lrn = res["fp64"][0], res["learned"][0]
        h_err = float(np.abs(res["learned"][1] - res["fp64"][1]).max() / np.abs(res["fp64"][1]).max())
        rows.append({"seed": sd, "cg_plain": sum(r[2] for r in base), "cg_learned": sum(r[2] for r in lrn),
                     "picard_plain": sum(r[1] for r in base), "picard_learned": sum(r[1] for r in lrn),
                     "s_plain": sum(r[0] for r in base[1:]), "s_learned": sum(r[0] for r in lrn[1:]),
                     "h_rel_diff": h_err})
        log(f"[gain] {nx}x{ny} seed {sd} ({precond}, every {every}): CG {rows[-1]['cg_plain']} -> {rows[-1]['cg_learned']}, "
            f"Picard {rows[-1]['picard_plain']} -> {rows[-1]['picard_learned']}, h diff {h_err:.1e}")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--glaciers", type=int, default=64)
    ap.add_argument("--val-glaciers", type=int, default=4)
    ap.add_argument("--per-glacier", type=int, default=320)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--loss", choices=["l2", "energy"], default="l2",
                    help="scale-free L2 loss, or the same in the energy norm of the scaled operator")
    ap.add_argument("--out", default=str(LN.DEFAULT_WEIGHTS))
    ap.add_argument("--gain-steps", type=int, default=4)
    ap.add_argument("--gain-seeds", default="5001,5002")
    ap.add_argument("--kind", choices=["picard", "cg"], default="picard",
                    help="pairs of Picard corrections (warm start) or of CG iterations' remaining errors "
                         "(learned preconditioner)")
    ap.add_argument("--precond", choices=["multigrid", "chebyshev"], default="multigrid")
    ap.add_argument("--every", type=int, default=16, help="CG pairs: one pair per this many CG iterations")
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    x, t, counts = LN.generate_pairs(args.glaciers, args.nx, args.ny, args.per_glacier, log=log, kind=args.kind,
                                     precond=args.precond, every=args.every)
    t_pairs = time.perf_counter() - t0
    ntr = args.glaciers - args.val_glaciers
    idx_tr = np.concatenate([np.arange(k * args.per_glacier, k * args.per_glacier + counts[k]) for k in range(ntr)])
    idx_va = np.concatenate([np.arange(k * args.per_glacier, k * args.per_glacier + counts[k])
                             for k in range(ntr, args.glaciers)])
    H8, W8 = LN.canvas_shape(args.nx, args.ny)
    net = LN.UNet(max_batch=args.batch, max_h=H8, max_w=W8, seed=args.seed, loss=args.loss)
    val0 = LN.evaluate(net, x, t, idx_va, args.batch)
    t1 = time.perf_counter()
    hist, val = LN.train(net, x, t, idx_tr, iters=args.iters, batch=args.batch, lr=args.lr, seed=args.seed,
                         idx_val=idx_va, log=log, log_every=max(1, args.iters // 10))
    t_train = time.perf_counter() - t1
    meta = dict(glaciers=args.glaciers, per_glacier=args.per_glacier, nx=args.nx, ny=args.ny, iters=args.iters,
                batch=args.batch, lr=args.lr, seed=args.seed, loss=args.loss, val_loss=val[-1][1], kind=args.kind,
                precond=args.precond, every=args.every)
    net.save(args.out, **meta)
    log(f"[train] saved {args.out}")
    rows = gain(args.out, args.nx, args.ny, [int(v) for v in args.gain_seeds.split(",")], args.gain_steps,
                args.precond, args.every if args.kind == "cg" else 0)
    print(json.dumps({
        "pairs": int(sum(counts)), "pairs_per_glacier": counts[:4], "pair_seconds": round(t_pairs, 1),
        "train_seconds": round(t_train, 1), "train_iters": args.iters, "batch": args.batch,
        "train_loss_first_last": [float(np.mean(hist[:50])), float(np.mean(hist[-50:]))],
        "val_loss_untrained": val0, "val_loss": val, f"gain_{args.nx}x{args.ny}_{args.precond}": rows,
    }), flush=True)


if __name__ == "__main__":
    main()
