"""The library's multigrid preconditioner (one V-cycle, shk_apply_precond)
against the NumPy restatement (tests/vcycle_restate.py) on the same assembled
system, on shapes that between them run every V-cycle kernel: two-column
red-black legs with FP64 and FP32 right-hand sides and with BF16 and FP32
coefficients, at two rows per wave and at 4 and 8 (2052 x 1023, 4100 x 1025:
the finest level's marching legs with their loads issued a step ahead, last
tiles of 3 and 1 rows), the flat legs on odd-width and small levels, and the
LDS tail with the dense coarsest solve.  FP32
inside: the two agree to 1e-5 of |z| (rounding order and the hardware
reciprocal); a wrong neighbour or smoothing order is off by orders of magnitude.
Symmetry r2'M r1 = r1'M r2 is checked on the library's own output."""
import ctypes as C

import numpy as np
import pytest

from helpers import hip_state, true_residuals
from vcycle_restate import VCycle, library_nu2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(130, 66), (520, 130), (1282, 150), (6, 6), (2052, 1023), (4100, 1025)])
def test_vcycle_matches_restatement(hip, shape):
    from shakti_hip import synthetic as S
    from shakti_hip import api
    from shakti_hip import _lib as L
    nx, ny = shape
    gg, raw = S.glacier(nx, ny, bed_amp=50.0, seed=11)
    hg, hp, hs = hip_state(hip, (gg.nx, gg.ny, gg.lx, gg.ly), raw)
    hs._bind(p=hp, kfs=hip.Harmonic())
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
    api._bind_solver(hs, ls)
    lib = L.load()
    L.check(lib.shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    arrs = [np.zeros((nx, ny), order="F") for _ in range(6)]
    L.check(lib.shk_download_system(hs.ctx, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]), hs.ctx)
    aP, aE, _, aN, _, _ = arrs
    # the library's default smoothing weight and coarse scales (linear transfers
    # below the finest level); BF16 finest-level coefficients where the finest
    # level takes the two-column legs
    two_col = nx % 2 == 0 and nx * ny > 4096  # the finest level takes the BF16 two-column legs
    M = VCycle(aP, aE, aN, bf16_level0=two_col, bil_from=1, scale_bil=0.8, nu2=library_nu2, nu2_fine=two_col)
    rng = np.random.default_rng(7)
    rs = [np.asfortranarray(rng.standard_normal((nx, ny))) for _ in range(2)]
    zs = []
    for r in rs:
        z = np.zeros((nx, ny), order="F")
        L.check(lib.shk_apply_precond(hs.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                      z.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
        zr = M(r)
        err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
        print(f"shape {shape}: |z - z_restated| / |z| = {err:.2e}", flush=True)
        assert err < 1e-5, err
        zs.append(z)
    a, b = float(np.sum(rs[1] * zs[0])), float(np.sum(rs[0] * zs[1]))
    scale = np.sqrt(float(np.sum(rs[0] * zs[0])) * float(np.sum(rs[1] * zs[1])))
    assert abs(a - b) <= 1e-4 * scale, (a, b)


def test_vcycle_parabolic_system(hip):
    """The backward-Euler system (storage term e_v / dt on the diagonal) takes the
    unscaled coarse correction (scale 1.0): the same comparison with the restatement."""
    from shakti_hip import synthetic as S
    from shakti_hip import api
    from shakti_hip import _lib as L
    nx, ny = 520, 130
    gg, raw = S.glacier(nx, ny, bed_amp=50.0, seed=12)
    hg, hp, hs = hip_state(hip, (gg.nx, gg.ny, gg.lx, gg.ly), raw, p_kw={"e_v": 1e-3})
    hs._bind(p=hp, kfs=hip.Harmonic())
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
    api._bind_solver(hs, ls)
    lib = L.load()
    L.check(lib.shk_assemble(hs.ctx, 1, 600.0), hs.ctx)
    arrs = [np.zeros((nx, ny), order="F") for _ in range(6)]
    L.check(lib.shk_download_system(hs.ctx, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]), hs.ctx)
    M = VCycle(arrs[0], arrs[1], arrs[3], omega=1.2, scale=1.0, bf16_level0=True, bil_from=1, scale_bil=1.0,
               nu2=library_nu2, nu2_fine=True)
    r = np.asfortranarray(np.random.default_rng(8).standard_normal((nx, ny)))
    z = np.zeros((nx, ny), order="F")
    L.check(lib.shk_apply_precond(hs.ctx, r.ctypes.data_as(C.POINTER(C.c_double)),
                                  z.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
    zr = M(r)
    err = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    print(f"parabolic {nx}x{ny}: |z - z_restated| / |z| = {err:.2e}", flush=True)
    assert err < 1e-5, err


def test_parabolic_solve_after_an_elliptic_one_matches_a_fresh_context(hip):
    """CG batches are replayed from captured HIP graphs, and the V-cycle's coarse
    scale (a kernel argument) differs between elliptic and parabolic systems: a
    parabolic multigrid solve after an elliptic one on the same context takes the
    same CG iterations and gives the same head, bit for bit, as on a fresh context."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(257, 66, bed_amp=50.0, seed=4)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for warm in (True, False):
        hg, hp, hs = hip_state(hip, g_args, raw, p_kw={"e_v": 1e-3})
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
        h0 = hs.get("h")
        if warm:
            hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
            hs.set("h", h0)
        hip.solve_parabolic_linear_system(ls, hs, hg, hp, hip.Harmonic(), 600.0)
        out.append((ls.last_stats.niter, hs.get("h")))
    print(f"parabolic CG iterations: after an elliptic solve {out[0][0]}, fresh {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("shape", [(2048, 1024), (8192, 2048)])
def test_fp32_z_gives_the_same_iterates(oracle, hip, monkeypatch, shape):
    """FP64 CG with the V-cycle keeps z in FP32 where the deferred finish runs, the
    finest level is BF16 and the matvec is the two-column kernel: the up-leg
    stores the FP32 values it computes and the matvec widens them on load, so
    the CG iterates are those of the FP64-z path (SHK_NO_DEFER=1, which
    finishes each reduction in its own kernel in the same order), bit for bit:
    the same iteration count and the same head.  At 8192 x 2048 the deferred
    matvec runs many more workgroups than the chip holds at once, so a
    workgroup that started after the stop decision of its own launch would
    show here (that decision reaches the next launch only).  The head is also
    held to its FP64 residual, recomputed on the host."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(*shape, bed_amp=50.0, seed=5)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for no_defer in (False, True):
        if no_defer:
            monkeypatch.setenv("SHK_NO_DEFER", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        h0 = hs.get("h")
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        out.append((ls.last_stats.niter, hs.get("h"), ls.last_stats.true_residual, h0, hs))
        monkeypatch.delenv("SHK_NO_DEFER", raising=False)
    print(f"CG iterations: FP32 z {out[0][0]}, FP64 z {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])
    rel_b, red = true_residuals(oracle, out[0][4], out[0][3])
    print(f"||b - A h|| / ||b|| = {rel_b:.3e} (library {out[0][2]:.3e}), reduction from the start {red:.3e}",
          flush=True)
    assert abs(out[0][2] - rel_b) <= 1e-3 * rel_b
    assert red <= 1e-8


def test_non_temporal_cg_matvec_gives_the_same_iterates(hip, monkeypatch):
    """At 4096 x 1024 a CG iteration's arrays exceed the Infinity Cache and the
    CG matvec of FP64 CG with the V-cycle loads its coefficients and stores p and
    Ap with non-temporal hints; the arithmetic is that of the cached kernel (the
    one the 2048 x 1024 parity case runs), so a context created with
    SHK_NO_NT=1 takes the same CG iterations and gives the same head, bit for bit."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(4096, 1024, bed_amp=50.0, seed=6)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for no_nt in (False, True):
        if no_nt:
            monkeypatch.setenv("SHK_NO_NT", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        out.append((ls.last_stats.niter, hs.get("h")))
        monkeypatch.delenv("SHK_NO_NT", raising=False)
    print(f"CG iterations: non-temporal {out[0][0]}, cached {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])


def test_fused_cg_update_gives_the_same_iterates(hip, monkeypatch):
    """FP64 CG with the V-cycle fuses r -= alpha Ap into the finest down-leg,
    which then writes the updated r to a second buffer (rows and columns shared
    with neighbouring tiles are read before any tile writes).  On a grid whose
    last tile is ragged (2052 x 1023: four rows per wave, three in the last
    tile, odd ny) the iterates are those of the separate update kernel
    (SHK_NO_FUSED_UPDATE=1), bit for bit."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(2052, 1023, bed_amp=50.0, seed=7)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for separate in (False, True):
        if separate:
            monkeypatch.setenv("SHK_NO_FUSED_UPDATE", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        out.append((ls.last_stats.niter, hs.get("h")))
        monkeypatch.delenv("SHK_NO_FUSED_UPDATE", raising=False)
    print(f"CG iterations: fused update {out[0][0]}, separate {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0] > 0
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
def test_captured_graphs_give_the_iterates_of_kernel_by_kernel_launches(hip, monkeypatch, precond):
    """CG batches are captured once per batch size into hipGraphs and replayed;
    with SHK_NO_GRAPHS=1 the same kernels are launched one by one.  Two Picard
    iterations (several batch sizes, graphs re-keyed between solves) give the
    same CG counts and the same head, bit for bit."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(1030, 514, bed_amp=50.0, seed=8)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for no_graphs in (False, True):
        if no_graphs:
            monkeypatch.setenv("SHK_NO_GRAPHS", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, check_every=16, **kw)
        ps = hip.PicardSolver(2, 1e-6, ls, hg)
        hip.elliptic_solver(ps, hs, hg, hp, hip.Harmonic(), hip.RegularizedCoulombSlidingLaw(0.5))
        out.append(([e["niter"] for e in hs.solve_log()], hs.get("h")))
        monkeypatch.delenv("SHK_NO_GRAPHS", raising=False)
    print(f"{precond}: CG iterations with graphs {out[0][0]}, without {out[1][0]}", flush=True)
    assert out[0][0] == out[1][0] and len(out[0][0]) == 2
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("itmax,check_every", [(0, 8), (0, 3), (0, 1), (7, 8), (6, 8)])
def test_paired_x_steps_give_the_iterates_of_single_steps(hip, monkeypatch, itmax, check_every):
    """FP64 CG with the V-cycle's FP32 z takes the x steps in pairs: the matvec
    of an odd iteration leaves its step owed and the next one takes both,
    x = (x + a1 p1) + a2 p2, the operations of two single steps.  Whatever
    iteration and batch the solve stops at (converged inside a batch or at its
    end, batches of odd and even size, or tired after an odd or an even count),
    the head is that of single steps (SHK_NO_PAIR_X=1), bit for bit."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(2052, 1023, bed_amp=50.0, seed=8)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = []
    for single in (False, True):
        if single:
            monkeypatch.setenv("SHK_NO_PAIR_X", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10,
                                   itmax=itmax, check_every=check_every)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        st = ls.last_stats
        out.append((st.niter, st.true_residual, hs.get("h")))
        monkeypatch.delenv("SHK_NO_PAIR_X", raising=False)
    print(f"itmax {itmax}, check_every {check_every}: CG iterations paired {out[0][0]}, single {out[1][0]}",
          flush=True)
    assert out[0][0] == out[1][0] > 0
    if itmax:
        assert out[0][0] == itmax
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
