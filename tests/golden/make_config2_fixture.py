"""Writes tests/golden/config2_1024x256_200steps.npz: the oracle's trajectory
of the 1024 x 256 configuration with direct solves.

Configuration: synthetic glacier seed 2, bed amplitude 10 m; harmonic
K_face; implicit gap; regularised-Coulomb C = 0.5; Picard 500 / 1e-6 with
the test every iteration and no relaxation; 200 steps of 3600 s; forcing
ieb = uniform(t) + 20 seeded moulins of 1-10 m^3/s, uniform(t) the
SeasonalMeltInput of the one-day 1e-8 m/s + 1e-7 m/s sinusoid
(src/melt_input.jl:66-94).

Every Picard iteration assembles the system with the oracle
(src/linear_solver.jl:438-523), solves it by LAPACK's band Cholesky
(dpbtrf/dpbtrs through scipy, the stand-in for CHOLMOD's cholesky!/ldiv! of
src/linear_solver.jl:709-739: the oracle's own band Cholesky gives the same
heads to 2e-13 but takes ten times longer here), then refreshes the
derived fields with the oracle (src/elliptic_solver.jl:187-208); the Picard
test and step order restate src/elliptic_solver.jl:138-176 and
src/run.jl:82-88 as the oracle's orc_picard_loop / orc_step do.

Stored: the Picard (converged, last_iter) of every step; per step max|h|,
sum h, max b and sum b; h and b after steps 1 and 200; a checksum of the
inputs.  About 30 minutes on 8 cores; with --checkpoint PATH the State and
the records are saved there every 10 steps and a rerun resumes from them:

    python tests/golden/make_config2_fixture.py [--checkpoint PATH]
"""
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "shakti.jl_amd"))

OUT = pathlib.Path(__file__).resolve().parent / "config2_1024x256_200steps.npz"
NX, NY, SEED, AMP, STEPS, DT = 1024, 256, 2, 10.0, 200, 3600.0
KEEP = (1, STEPS)


def band_cholesky_solve(co):
    """The SPD system of the assembled coefficients by LAPACK's band Cholesky,
    unknowns renumbered iy-fastest (half-bandwidth ny)."""
    import scipy.linalg as sl
    aP, aE, aW, aN, aS, rhs = co
    nx, ny = aP.shape
    u = ny
    ab = np.zeros((u + 1, nx * ny))
    ab[u] = aP.ravel(order="C")
    s1 = np.zeros((nx, ny))
    s1[:, 1:] = -aS[:, 1:]  # coupling to (ix, iy-1)
    ab[u - 1] = s1.ravel(order="C")
    sn = np.zeros((nx, ny))
    sn[1:, :] = -aW[1:, :]  # coupling to (ix-1, iy)
    ab[0] = sn.ravel(order="C")
    c = sl.cholesky_banded(ab, lower=False, check_finite=True)
    x = sl.cho_solve_banded((c, False), rhs.ravel(order="C"), check_finite=False)
    return x.reshape((nx, ny), order="C")


def inputs():
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(NX, NY, bed_amp=AMP, seed=SEED)
    moulin = S.moulin_field(gg, raw["mask"], n=20, seed=SEED)
    return gg, raw, moulin


def trajectory(log=print, checkpoint=None):
    import oracle as O
    from shakti_hip import synthetic as S
    gg, raw, moulin = inputs()
    g = O.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = O.Model(kfs=1, gap_scheme=0, sliding_C=0.5)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    eps = np.finfo(np.float64).eps
    out = {"zb_sum": np.float64(np.sum(raw["zb"])), "moulin_sum": np.float64(np.sum(moulin))}
    picard, stats = [], []
    first = 1
    if checkpoint and pathlib.Path(checkpoint).exists():
        with np.load(checkpoint, allow_pickle=False) as c:
            for f in O.ALL_FIELDS:
                getattr(s, f)[...] = c["state_" + f]
            picard = [tuple(r) for r in c["picard"]]
            stats = [tuple(r) for r in c["stats"]]
            out.update({k: c[k] for k in c.files if k[:2] in ("h_", "b_")})
        first = len(picard) + 1
        log(f"resumed after step {first - 1}", flush=True)
    t = (first - 1) * DT
    for step in range(first, STEPS + 1):
        t0 = time.time()
        s.ieb[...] = O.seasonal_ieb(t, **S.SEASONAL_C2) + moulin  # update_ieb!, then the moulins on top
        conv, last = False, 500
        for it in range(1, 501):
            h_prev = s.h.copy()
            s.h[...] = band_cholesky_solve(O.assemble_mfls(s, g, p, 1))
            O.update_derived(s, g, p, m)
            d, a = np.abs(s.h - h_prev), np.abs(s.h)
            if not (np.isnan(d).any() or np.isnan(a).any()) and d.max() / (a.max() + eps) < 1e-6:
                conv, last = True, it
                break
        O.step_b(s, g, p, m, DT)
        t += DT
        picard.append((int(conv), int(last)))
        stats.append((np.max(np.abs(s.h)), np.sum(s.h), np.max(s.b), np.sum(s.b)))
        if step in KEEP:
            out[f"h_{step}"] = s.h.copy()
            out[f"b_{step}"] = s.b.copy()
        log(f"step {step}: Picard {last} (converged {conv}), {time.time() - t0:.1f} s", flush=True)
        if checkpoint and step % 10 == 0:
            np.savez(checkpoint, picard=np.array(picard, dtype=np.int64), stats=np.array(stats),
                     **{"state_" + f: getattr(s, f) for f in O.ALL_FIELDS},
                     **{k: v for k, v in out.items() if k[:2] in ("h_", "b_")})
    out["picard"] = np.array(picard, dtype=np.int64)
    out["stats"] = np.array(stats, dtype=np.float64)
    return out


if __name__ == "__main__":
    import argparse
    import oracle as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--checkpoint", default=None)
    O.build()
    np.savez_compressed(OUT, **trajectory(checkpoint=ap.parse_args().checkpoint))
    print(f"wrote {OUT}")
