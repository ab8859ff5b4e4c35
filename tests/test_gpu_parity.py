"""HIP library vs the CPU oracle on identical inputs (runs on an MI355X).

Tolerances: masks and masked face fields bit-exact; FP64 fields after the
initial conditions and derived-field refresh within 1e-13 of the field's
maximum (pow() may differ by an ulp); heads after a solve within 1e-9
relative and gap heights within 1e-8 relative, per step (the stated bar)."""
import numpy as np
import pytest

from helpers import FIELDS_EXACT, field_errors, hip_state, oracle_state, rel_err

pytestmark = pytest.mark.gpu

ALL = None


def _fixture(O):
    g, p, m, raw = O.test_fixture_6x6()
    return (g.nx, g.ny, g.lx, g.ly), raw


def test_initial_conditions_6x6(oracle, hip):
    g_args, raw = _fixture(oracle)
    _, _, _, os_ = oracle_state(oracle, g_args, raw, C=0.25)
    _, _, hs = hip_state(hip, g_args, raw, C=0.25)
    from shakti_hip._lib import FIELD_NAMES
    for n in FIELDS_EXACT:
        assert np.array_equal(hs.get(n), getattr(os_, n)), n
    errs = field_errors(hs, os_, [n for n in FIELD_NAMES if n not in FIELDS_EXACT])
    bad = {k: v for k, v in errs.items() if v > 1e-13}
    assert not bad, bad


def test_initial_conditions_synthetic(oracle, hip):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(96, 40, bed_amp=10.0, seed=3)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    _, _, _, os_ = oracle_state(oracle, g_args, raw)
    _, _, hs = hip_state(hip, g_args, raw)
    from shakti_hip._lib import FIELD_NAMES
    for n in FIELDS_EXACT:
        assert np.array_equal(hs.get(n), getattr(os_, n)), n
    errs = field_errors(hs, os_, [n for n in FIELD_NAMES if n not in FIELDS_EXACT])
    bad = {k: v for k, v in errs.items() if v > 1e-13}
    assert not bad, bad


@pytest.mark.parametrize("kfs", [0, 1])
@pytest.mark.parametrize("parabolic", [False, True])
def test_assembly_matches_oracle(oracle, hip, kfs, parabolic):
    """Perturbed mid-solve state (h += 0.01*(1:n)), derived refresh, then the
    matrix-free assembly: all five coefficient arrays and rhs."""
    g_args, raw = _fixture(oracle)
    p_kw = {"e_v": 1e-3} if parabolic else {}
    g, p, m, os_ = oracle_state(oracle, g_args, raw, p_kw=p_kw, kfs=kfs, C=0.25)
    hg, hp, hs = hip_state(hip, g_args, raw, p_kw=p_kw, C=0.25)
    dh = 0.01 * np.arange(1, g.nx * g.ny + 1, dtype=np.float64).reshape((g.nx, g.ny), order="F")
    os_.h += dh
    oracle.update_derived(os_, g, p, m)
    hs.set("h", hs.get("h") + dh)
    hip.update_derived(hs, hp, hip.RegularizedCoulombSlidingLaw(0.25))
    from shakti_hip._lib import FIELD_NAMES
    errs = field_errors(hs, os_, [n for n in FIELD_NAMES if n not in FIELDS_EXACT])
    assert max(errs.values()) < 1e-13, errs
    co = oracle.assemble_mfls(os_, g, p, kfs, parabolic, 60.0)
    import ctypes as C
    from shakti_hip import _lib as L
    hs._bind(p=hp, kfs=[hip.Arithmetic(), hip.Harmonic()][kfs])
    L.check(L.load().shk_assemble(hs.ctx, int(parabolic), 60.0), hs.ctx, "assemble")
    arrs = [np.zeros((g.nx, g.ny), order="F") for _ in range(6)]
    L.check(L.load().shk_download_system(hs.ctx, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]), hs.ctx)
    for name, a, b in zip(["aP", "aE", "aW", "aN", "aS", "rhs"], arrs, co):
        assert rel_err(a, b) < 1e-14, (name, rel_err(a, b))


def test_operator_apply_matches_oracle(oracle, hip):
    """y = A x for a random x on a synthetic 300x70 glacier (ragged against
    the 64-wide strips), against the oracle's 5-array matvec."""
    from shakti_hip import synthetic as S
    import ctypes as C
    from shakti_hip import _lib as L
    gg, raw = S.glacier(300, 70, bed_amp=10.0, seed=5)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    hg, hp, hs = hip_state(hip, g_args, raw)
    co = oracle.assemble_mfls(os_, g, p, 1)
    hs._bind(p=hp, kfs=hip.Harmonic())
    L.check(L.load().shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    rng = np.random.default_rng(0)
    x = np.asfortranarray(rng.standard_normal((g.nx, g.ny)))
    y = np.zeros_like(x)
    L.check(L.load().shk_apply_operator(hs.ctx, x.ctypes.data_as(C.POINTER(C.c_double)),
                                        y.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
    y_ref = oracle.matvec(co, x)
    assert rel_err(y, y_ref) < 1e-15


def test_residual_norms_match_oracle(oracle, hip):
    """shk_residual_norms: ||b - A x|| and ||b|| of the last assembly, against
    the oracle's five-array operator on the library's own assembly, for a
    random x and for the Cholesky solution (the residual at the FP64 floor)."""
    from shakti_hip import synthetic as S
    import ctypes as C
    from shakti_hip import _lib as L
    gg, raw = S.glacier(300, 70, bed_amp=10.0, seed=5)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    g, p, m, os_ = oracle_state(oracle, g_args, raw)
    hg, hp, hs = hip_state(hip, g_args, raw)
    hs._bind(p=hp, kfs=hip.Harmonic())
    L.check(L.load().shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    sysd = hs.download_system()
    co = [sysd[k] for k in ("aP", "aE", "aW", "aN", "aS")]
    b = sysd["rhs"]
    x_chol = oracle.cholesky_solve(co + [b], b)
    for x in (np.asfortranarray(np.random.default_rng(2).standard_normal(b.shape)), x_chol):
        out = np.zeros(2)
        L.check(L.load().shk_residual_norms(hs.ctx, x.ctypes.data_as(C.POINTER(C.c_double)),
                                            out.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
        r = b - oracle.matvec(co, x)
        assert out[1] == pytest.approx(np.linalg.norm(b), rel=1e-13)
        assert out[0] == pytest.approx(np.linalg.norm(r), rel=1e-6, abs=1e-14 * np.linalg.norm(b))


@pytest.mark.parametrize("precond", ["jacobi", "chebyshev"])
def test_single_solve_vs_cholesky_6x6(oracle, hip, precond):
    g_args, raw = _fixture(oracle)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, C=0.25)
    hg, hp, hs = hip_state(hip, g_args, raw, C=0.25)
    sv = oracle.Solver(g, "cholesky")
    sv.solve_elliptic(os_, p, 0)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem,
                               chebyshev_degree=4 if precond == "chebyshev" else None, atol=0.0, rtol=1e-14)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Arithmetic())
    assert ls.ws.stats.niter > 0
    assert rel_err(hs.get("h"), os_.h) < 1e-12


@pytest.mark.parametrize("bounds_mode", ["reference", "gershgorin"])
def test_chebyshev_bounds_and_iterations_match_oracle_cg(oracle, hip, bounds_mode):
    """Same Lanczos bounds (to rounding) and the same CG iteration count as the
    oracle's restated Krylov CG with the Chebyshev preconditioner, with the
    reference's bounds and with the Gershgorin lambda_max."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(128, 48, bed_amp=10.0, seed=2)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    hg, hp, hs = hip_state(hip, g_args, raw)
    co = oracle.assemble_mfls(os_, g, p, 1)
    gersh = bounds_mode == "gershgorin"
    x_ref, st_ref, bounds = oracle.cg_solve(co, co[5], os_.h, precond="chebyshev", atol=0.0, rtol=1e-10,
                                            gershgorin=gersh)
    est = (oracle.estimate_eigenvalue_bounds_d if gersh else oracle.estimate_eigenvalue_bounds)(co, co[5])
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-10,
                               check_every=1, chebyshev_bounds=bounds_mode)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
    st = ls.last_stats
    assert abs(ls.precond.lambda_min - bounds[0]) <= 1e-6 * abs(bounds[0])
    assert abs(ls.precond.lambda_max - bounds[1]) <= 1e-6 * abs(bounds[1])
    assert abs(st.lz_lambda_min - est[0]) <= 1e-6 * abs(est[0])
    assert abs(st.lz_lambda_max - est[1]) <= 1e-6 * abs(est[1])
    if gersh:
        # the Gershgorin bound is a max of per-row sums: the same bits as the oracle's
        assert st.gershgorin == oracle.gershgorin(co) == bounds[1]
    assert abs(ls.ws.stats.niter - st_ref.niter) <= 1, (ls.ws.stats.niter, st_ref.niter)
    assert rel_err(hs.get("h"), x_ref) < 1e-9
    # the FP64 residual of the returned head, recomputed by the oracle's operator
    r = co[5] - oracle.matvec(co, hs.get("h"))
    tr = np.linalg.norm(r) / np.linalg.norm(co[5])
    assert abs(ls.ws.stats.true_residual - tr) <= 1e-6 * tr + 1e-300


def test_picard_loop_6x6_matches_cholesky_oracle(oracle, hip):
    g_args, raw = _fixture(oracle)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, C=0.25)
    hg, hp, hs = hip_state(hip, g_args, raw, C=0.25)
    sv = oracle.Solver(g, "cholesky")
    conv, last, _ = oracle.picard_loop(sv, os_, g, p, m, 500, 1e-6)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, hg)
    hip.elliptic_solver(ps, hs, hg, hp, hip.Arithmetic(), hip.RegularizedCoulombSlidingLaw(0.25))
    assert ps.converged == conv and ps.last_iter == last
    assert rel_err(hs.get("h"), os_.h) < 1e-9
    for n in ["N", "mdot", "K", "q_x", "Re"]:
        assert rel_err(hs.get(n), getattr(os_, n)) < 1e-8, n


def test_config1_trajectory_48_steps(oracle, hip):
    """64x32, Cholesky oracle vs HIP Chebyshev(4)-PCG, 48 hourly steps: per
    step head <= 1e-9 and gap height <= 1e-8 relative, identical Picard
    counts, masks bit-exact."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(64, 32, bed_amp=10.0, seed=1)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    g, p, m, os_ = oracle_state(oracle, g_args, raw)
    hg, hp, hs = hip_state(hip, g_args, raw)
    sv = oracle.Solver(g, "cholesky")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, hg)
    sim = hip.Simulation(hg, hs, 48, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         hip.RegularizedCoulombSlidingLaw(0.5), ps=ps)
    worst_h = worst_b = 0.0
    for t in range(48):
        conv, last, _ = oracle.step(sv, os_, g, p, m, 3600.0)
        hip.step(sim)
        sim.total_time += sim.dt
        assert ps.converged == conv and ps.last_iter == last, (t, ps.last_iter, last)
        eh, eb = rel_err(hs.get("h"), os_.h), rel_err(hs.get("b"), os_.b)
        worst_h, worst_b = max(worst_h, eh), max(worst_b, eb)
        assert eh <= 1e-9 and eb <= 1e-8, (t, eh, eb)
    for n in FIELDS_EXACT:
        assert np.array_equal(hs.get(n), getattr(os_, n)), n
    print(f"config-1 worst per-step errors: h {worst_h:.2e}  b {worst_b:.2e}")


def test_parabolic_steps_match_oracle(oracle, hip):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(80, 24, bed_amp=20.0, seed=4)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    p_kw = {"e_v": 1e-3}
    g, p, m, os_ = oracle_state(oracle, g_args, raw, p_kw=p_kw)
    hg, hp, hs = hip_state(hip, g_args, raw, p_kw=p_kw)
    sv = oracle.Solver(g, "cholesky")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    sim = hip.Simulation(hg, hs, 10, 600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         hip.RegularizedCoulombSlidingLaw(0.5), ls=ls)
    for t in range(10):
        oracle.step(sv, os_, g, p, m, 600.0)
        hip.step(sim)
        assert rel_err(hs.get("h"), os_.h) <= 1e-9, t
        assert rel_err(hs.get("b"), os_.b) <= 1e-8, t


@pytest.mark.parametrize("degree", [2, 3, 4, 8])
def test_fused_update_chebyshev_matches_unfused(oracle, hip, degree, monkeypatch):
    """The fused update + Chebyshev kernel gives the same solve as the
    separate update / sweep kernels (same arithmetic; only the dot-product
    partition differs), on a grid ragged against the strip widths."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(301, 77, bed_amp=20.0, seed=6)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    out = {}
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("SHK_NO_FUSED", raising=False)
        else:
            monkeypatch.setenv("SHK_NO_FUSED", "1")
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=degree, atol=0.0,
                                   rtol=1e-12, check_every=2)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        out[fused] = (hs.get("h"), ls.ws.stats.niter)
    assert abs(out[True][1] - out[False][1]) <= 2, (out[True][1], out[False][1])
    assert rel_err(out[True][0], out[False][0]) < 1e-11
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    co = oracle.assemble_mfls(os_, g, p, 1)
    x_ref, st_ref, _ = oracle.cg_solve(co, co[5], os_.h, precond="chebyshev", degree=degree, atol=0.0, rtol=1e-12)
    assert rel_err(out[True][0], x_ref) < 1e-10


@pytest.mark.parametrize("kfs", [0, 1])
def test_fp32_refinement_solve_matches_fp64(oracle, hip, kfs):
    """The FP32-inner / FP64-refinement form reaches the same FP64 residual
    test as the FP64 form: heads agree to the solve tolerance."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(257, 65, bed_amp=30.0, seed=8)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    kf = [hip.Arithmetic(), hip.Harmonic()][kfs]
    res = {}
    for form in ("fp64", "fp32_refine"):
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-11,
                                   form=form, inner_rtol=1e-4)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, kf)
        res[form] = (hs.get("h"), ls.ws.stats)
    st = res["fp32_refine"][1]
    assert st.solved and st.refinements >= 2
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=kfs)
    co = oracle.assemble_mfls(os_, g, p, kfs)
    x_ref = oracle.cholesky_solve(co, co[5]) if g.nx * g.ny < 20000 else None
    e = rel_err(res["fp32_refine"][0], res["fp64"][0])
    assert e < 1e-9, e
    if x_ref is not None:
        assert rel_err(res["fp32_refine"][0], x_ref) < 1e-9


def test_fp32_refinement_picard_trajectory(oracle, hip):
    """Config-1 style run with the precision form: Picard converges every step
    and h, b stay within 1e-7 / 1e-6 of the Cholesky oracle."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(64, 32, bed_amp=10.0, seed=1)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    g, p, m, os_ = oracle_state(oracle, g_args, raw)
    hg, hp, hs = hip_state(hip, g_args, raw)
    sv = oracle.Solver(g, "cholesky")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-8,
                               form="fp32_refine", inner_rtol=1e-3)
    ps = hip.PicardSolver(500, 1e-6, ls, hg)
    sim = hip.Simulation(hg, hs, 12, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         hip.RegularizedCoulombSlidingLaw(0.5), ps=ps)
    for t in range(12):
        oracle.step(sv, os_, g, p, m, 3600.0)
        hip.step(sim)
        assert ps.converged
        assert rel_err(hs.get("h"), os_.h) <= 1e-7, t
        assert rel_err(hs.get("b"), os_.b) <= 1e-6, t


@pytest.mark.parametrize("shape", [(6, 6), (257, 65), (301, 77), (1283, 1001), (2048, 1024)])
def test_multigrid_fp32_refinement_solve(oracle, hip, shape):
    """FP32 multigrid-preconditioned CG inside FP64 refinement, and FP64 CG
    preconditioned by the FP32 V-cycle, solve to within 1e-9 of the FP64
    Chebyshev form in far fewer iterations.  The first four grids are ragged
    against the 2x2 coarsening (odd widths take the flat red-black legs);
    2048x1024 (2 M cells, even) runs the two-column legs with BF16 finest-level
    coefficients and the FP64 CG with the two-column matvec and the V-cycle's
    z kept in FP32."""
    from shakti_hip import synthetic as S
    nx, ny = shape
    if shape == (6, 6):
        g0, p0, m0, raw = oracle.test_fixture_6x6()
        g_args = (6, 6, g0.lx, g0.ly)
    else:
        gg, raw = S.glacier(nx, ny, bed_amp=30.0, seed=9)
        g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    res = {}
    for name, kw in [("fp64", dict(chebyshev_degree=4)),
                     ("mg", dict(form="fp32_refine", multigrid=True, inner_rtol=1e-4)),
                     ("fp64_mg", dict(form="fp64", multigrid=True))]:
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, **kw)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        res[name] = (hs.get("h"), ls.ws.stats)
    print(f"{shape}: CG iterations " + ", ".join(f"{k} {v[1].niter}" for k, v in res.items()), flush=True)
    for name in ("mg", "fp64_mg"):
        st = res[name][1]
        assert st.solved, name
        assert rel_err(res[name][0], res["fp64"][0]) < 1e-9, name
        if shape != (6, 6):
            assert st.niter < res["fp64"][1].niter / 3, (name, st.niter, res["fp64"][1].niter)


def test_streaming_kernels_on_a_grid_beyond_the_infinity_cache(hip):
    """At 4096x2048 (8.4 M cells, 336 MB per operator apply) the FP64 operator
    and CG kernels switch to the two-columns-per-lane, non-temporal variants.
    The operator apply equals the reference's 5-array matvec bit for bit
    (same products, same order), and the plain Chebyshev form (fused
    non-temporal kernel) and the multigrid form reach the same head."""
    from shakti_hip import synthetic as S
    import ctypes as C
    from shakti_hip import _lib as L
    gg, raw = S.glacier(4096, 2048, bed_amp=10.0, seed=12)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    print("4096x2048: operator apply", flush=True)
    hg, hp, hs = hip_state(hip, g_args, raw)
    hs._bind(p=hp, kfs=hip.Harmonic())
    L.check(L.load().shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    n = (gg.nx, gg.ny)
    arrs = [np.zeros(n, order="F") for _ in range(6)]
    L.check(L.load().shk_download_system(hs.ctx, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]), hs.ctx)
    aP, aE, aW, aN, aS, _ = arrs
    rng = np.random.default_rng(3)
    x = np.asfortranarray(rng.standard_normal(n))
    y = np.zeros_like(x)
    L.check(L.load().shk_apply_operator(hs.ctx, x.ctypes.data_as(C.POINTER(C.c_double)),
                                        y.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
    ref = aP * x
    ref[:-1] -= aE[:-1] * x[1:]
    ref[1:] -= aW[1:] * x[:-1]
    ref[:, :-1] -= aN[:, :-1] * x[:, 1:]
    ref[:, 1:] -= aS[:, 1:] * x[:, :-1]
    assert np.array_equal(y, ref)
    del hs
    res = {}
    for name, kw in [("cheb", dict(chebyshev_degree=4)), ("mg", dict(multigrid=True))]:
        print(f"4096x2048: solve {name}", flush=True)
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, check_every=32, **kw)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        assert ls.ws.stats.solved
        res[name] = hs.get("h")
        del hs
    assert rel_err(res["cheb"], res["mg"]) < 1e-9


@pytest.mark.parametrize("bounds_mode", ["reference", "gershgorin"])
def test_invalid_estimate_keeps_previous_bounds(oracle, hip, bounds_mode):
    """test/linear_solver_test.jl:197-222 through the C ABI: an all-zero rhs
    makes the estimate 0/0, so update_chebyshev_bounds keeps the placeholder
    (1, 1) with a warning; a real rhs then gives the oracle's bounds; a second
    zero rhs keeps those (the previous valid bounds), again with a warning."""
    import warnings
    from shakti_hip import _lib as L
    g_args, raw = _fixture(oracle)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, C=0.25)
    hg, hp, hs = hip_state(hip, g_args, raw, C=0.25)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, chebyshev_bounds=bounds_mode)
    hs._bind(p=hp, kfs=hip.Arithmetic())
    L.check(L.load().shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    zeros = np.zeros((g.nx, g.ny))
    with pytest.warns(UserWarning, match="invalid"):
        pc = hip.update_chebyshev_bounds(ls, hs, zeros)
    assert (pc.lambda_min, pc.lambda_max) == (1.0, 1.0)
    co = oracle.assemble_mfls(os_, g, p, 0)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        pc = hip.update_chebyshev_bounds(ls, hs, co[5])
    if bounds_mode == "reference":
        want = oracle.update_chebyshev_bounds(co, co[5])
    else:
        lo = oracle.estimate_eigenvalue_bounds_d(co, co[5])[0]
        want = (lo, oracle.gershgorin(co), False)
    assert not want[2]
    assert pc.lambda_min == pytest.approx(want[0], rel=1e-10)
    assert pc.lambda_max == pytest.approx(want[1], rel=1e-10)
    kept = (pc.lambda_min, pc.lambda_max)
    with pytest.warns(UserWarning, match="invalid"):
        pc = hip.update_chebyshev_bounds(ls, hs, zeros)
    assert (pc.lambda_min, pc.lambda_max) == kept
    assert np.isfinite(kept).all() and 0 < kept[0] < kept[1]


@pytest.mark.parametrize("case", ["explicit_gap", "relaxation", "seasonal", "seasonal_moulins"])
def test_trajectory_options_match_oracle(oracle, hip, case):
    """Six hourly steps at 64 x 32 against the oracle's Cholesky trajectory
    with each option the configurations use beyond config 1: the explicit gap
    scheme (src/gap_height.jl:127-135), head relaxation alpha = 0.7
    (src/elliptic_solver.jl:33-46), SeasonalMeltInput (ieb mode 1,
    src/melt_input.jl:86-94) and the seasonal input with moulins on top (ieb
    mode 2).  Same Picard counts, h within 1e-9 and b within 1e-8 each step."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(64, 32, bed_amp=10.0, seed=4)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    gap = 1 if case == "explicit_gap" else 0
    alpha = 0.7 if case == "relaxation" else None
    g, p, m, os_ = oracle_state(oracle, g_args, raw, gap=gap)
    hg, hp, hs = hip_state(hip, g_args, raw)
    moulin = S.moulin_field(gg, raw["mask"], n=5, seed=4) if case == "seasonal_moulins" else None
    if case == "seasonal":
        mi = hip.SeasonalMeltInput(**S.SEASONAL_C2)
    elif case == "seasonal_moulins":
        mi = hip.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
    else:
        mi = hip.ConstantMeltInput()
    sv = oracle.Solver(g, "cholesky")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, hg, alpha=alpha)
    sim = hip.Simulation(hg, hs, 6, 3600.0, hp, "explicit" if gap else "implicit", [], mi,
                         hip.RegularizedCoulombSlidingLaw(0.5), ps=ps)
    mode = {"seasonal": 1, "seasonal_moulins": 2}.get(case, 0)
    for t in range(6):
        u = oracle.seasonal_ieb(sim.total_time, **S.SEASONAL_C2) if mode else 0.0
        conv, last, _ = oracle.step(sv, os_, g, p, m, 3600.0, ieb_mode=mode, ieb_uniform=u, moulin=moulin,
                                    alpha=alpha)
        hip.step(sim)
        sim.total_time += sim.dt
        assert (ps.converged, ps.last_iter) == (conv, last), (t, ps.last_iter, last)
        assert rel_err(hs.get("h"), os_.h) <= 1e-9, t
        assert rel_err(hs.get("b"), os_.b) <= 1e-8, t
        if mode:
            assert rel_err(hs.get("ieb"), os_.ieb) <= 1e-15, t


_FIELD_LAWS = {
    "linear": lambda H, g, c: H.LinearSlidingLaw(g, c),
    "regularized_coulomb_field": lambda H, g, c: H.RegularizedCoulombFieldSlidingLaw(g, c),
    "regularized_coulomb_v0": lambda H, g, c: H.RegularizedCoulombV0SlidingLaw(g, c, 3e-7),
}


@pytest.mark.parametrize("law", list(_FIELD_LAWS))
def test_field_sliding_laws_match_oracle(oracle, hip, law):
    """LinearSlidingLaw, RegularizedCoulombFieldSlidingLaw and
    RegularizedCoulombV0SlidingLaw (src/sliding_law.jl:61-90, 270-301,
    422-452) with a random cell-centred C: the initial conditions match the
    oracle's field by field (taub within 1e-13 of its maximum, the face
    staggering of C done on the device), then four hourly steps match the
    oracle's Cholesky trajectory (same Picard counts, h within 1e-9, b within
    1e-8, taub within 1e-8)."""
    from shakti_hip import synthetic as S
    from shakti_hip._lib import FIELD_NAMES
    gg, raw = S.glacier(96, 40, bed_amp=10.0, seed=5)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    Cc = np.asfortranarray(np.random.default_rng(7).uniform(0.2, 0.8, (gg.nx, gg.ny)))
    if law == "linear":
        Cc = Cc * 1e2  # taub = C^2 N ub: C ~ 50 puts it in the range of the other laws
    g, p, m, os_ = oracle_state(oracle, g_args, raw,
                                model=lambda og: oracle.Model.field_law(og, law, Cc, v0=3e-7))
    hg, hp, hs = hip_state(hip, g_args, raw, law=lambda hg_: _FIELD_LAWS[law](hip, hg_, Cc))
    for n in FIELDS_EXACT:
        assert np.array_equal(hs.get(n), getattr(os_, n)), n
    errs = field_errors(hs, os_, [n for n in FIELD_NAMES if n not in FIELDS_EXACT])
    bad = {k: v for k, v in errs.items() if v > 1e-13}
    assert not bad, bad
    assert np.count_nonzero(os_.taub_x) > 1000
    sv = oracle.Solver(g, "cholesky")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, hg)
    sim = hip.Simulation(hg, hs, 4, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         _FIELD_LAWS[law](hip, hg, Cc), ps=ps)
    for t in range(4):
        conv, last, _ = oracle.step(sv, os_, g, p, m, 3600.0)
        hip.step(sim)
        sim.total_time += sim.dt
        assert (ps.converged, ps.last_iter) == (conv, last), (t, ps.last_iter, last)
        assert rel_err(hs.get("h"), os_.h) <= 1e-9, t
        assert rel_err(hs.get("b"), os_.b) <= 1e-8, t
        assert rel_err(hs.get("taub_x"), os_.taub_x) <= 1e-8, t
        assert rel_err(hs.get("taub_y"), os_.taub_y) <= 1e-8, t


def test_field_sliding_law_without_coefficient_is_refused(hip):
    """A field law set through shk_set_model with no coefficient uploaded:
    the derived-field refresh fails with SHK_ERR_STATE and leaves taub as it
    was."""
    import ctypes as C
    from shakti_hip import _lib as L
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(32, 16, bed_amp=10.0, seed=1)
    hg, hp, hs = hip_state(hip, (gg.nx, gg.ny, gg.lx, gg.ly), raw)
    before = hs.get("taub_x")
    lib = L.load()
    m = L.ShkModel(0, 0, 2, 0.5, 0.0)
    assert lib.shk_set_model(hs.ctx, C.byref(m)) == 0
    rc = lib.shk_update_derived(hs.ctx)
    assert rc == -4, rc  # SHK_ERR_STATE
    assert b"coefficient" in lib.shk_last_error(hs.ctx)
    assert np.array_equal(hs.get("taub_x"), before)


@pytest.mark.parametrize("form", ["fp64_mg", "fp32_mg"])
def test_picard_atol_trajectory_matches_oracle(oracle, hip, form):
    """SHK_CFG_PICARD_ATOL (later Picard solves held to rtol x the first
    solve's initial residual) at 96 x 40 with seasonal forcing and moulins:
    the Picard counts of the oracle's Cholesky trajectory every step, h within
    1e-9 and b within 1e-8, fewer CG iterations than without the flag, and an
    unknown flag bit refused."""
    import ctypes as C
    from shakti_hip import _lib as L
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(96, 40, bed_amp=10.0, seed=7)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    moulin = S.moulin_field(gg, raw["mask"], n=5, seed=7)
    kw = (dict(multigrid=True, rtol=1e-10, form="fp64") if form == "fp64_mg" else
          dict(multigrid=True, rtol=1e-8, form="fp32_refine", inner_rtol=1e-4))
    iters = {}
    for flag in (False, True):
        g, p, m, os_ = oracle_state(oracle, g_args, raw)
        hg, hp, hs = hip_state(hip, g_args, raw)
        sv = oracle.Solver(g, "cholesky")
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, picard_atol=flag, **kw)
        ps = hip.PicardSolver(500, 1e-6, ls, hg)
        mi = hip.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
        sim = hip.Simulation(hg, hs, 6, 3600.0, hp, "implicit", [], mi, hip.RegularizedCoulombSlidingLaw(0.5), ps=ps)
        total = 0
        for t in range(6):
            u = oracle.seasonal_ieb(sim.total_time, **S.SEASONAL_C2)
            conv, last, _ = oracle.step(sv, os_, g, p, m, 3600.0, ieb_mode=2, ieb_uniform=u, moulin=moulin)
            hip.step(sim)
            sim.total_time += sim.dt
            total += ps.cg_iters
            assert (ps.converged, ps.last_iter) == (conv, last), (flag, t)
            assert rel_err(hs.get("h"), os_.h) <= 1e-9, (flag, t)
            assert rel_err(hs.get("b"), os_.b) <= 1e-8, (flag, t)
        iters[flag] = total
    print(f"{form}: CG iterations over 6 steps {iters[False]} without, {iters[True]} with the flag", flush=True)
    assert iters[True] < iters[False]
    cfg = ls.cfg()
    cfg.flags = 2
    assert L.load().shk_set_solver_cfg(hs.ctx, C.byref(cfg)) == -1  # SHK_ERR_ARG


_REFRESH_LAWS = {
    "coulomb": lambda H, g, c: H.RegularizedCoulombSlidingLaw(0.5),
    "prescribed": lambda H, g, c: H.PrescribedSlidingLaw(),
    "linear_field": lambda H, g, c: H.LinearSlidingLaw(g, c * 1e2),
    "coulomb_v0_field": lambda H, g, c: H.RegularizedCoulombV0SlidingLaw(g, c, 3e-7),
}


@pytest.mark.parametrize("shape", [(1283, 1001), (126, 65)])
@pytest.mark.parametrize("case", list(_REFRESH_LAWS) + ["no_sensible"])
def test_one_pass_refresh_matches_two_pass_kernels(hip, monkeypatch, shape, case):
    """The one-pass derived-field refresh (63 owned columns per wave, bands of
    rows) and the lean refresh of the Picard iterations that another one
    follows (N, K and mdot only) leave the State of the two-pass kernels bit
    for bit: after the initial conditions and after each of three hourly
    steps of the multigrid form (same Picard counts).  1283 x 1001 has a
    partial last strip and a one-row last band; 126 = 2 x 63 puts the east
    boundary face on the last strip's halo lane.  Each sliding law, and the
    melt rate without the sensible-heat term (ct = 0)."""
    from shakti_hip import synthetic as S
    from shakti_hip._lib import FIELD_NAMES
    nx, ny = shape
    gg, raw = S.glacier(nx, ny, bed_amp=10.0, seed=21)
    g_args = (gg.nx, gg.ny, gg.lx, gg.ly)
    Cc = np.asfortranarray(np.random.default_rng(8).uniform(0.2, 0.8, (nx, ny)))
    law = _REFRESH_LAWS.get(case, _REFRESH_LAWS["coulomb"])
    p_kw = {"ct": 0.0} if case == "no_sensible" else None
    runs = {}
    for two_pass in (True, False):
        if two_pass:
            monkeypatch.setenv("SHK_DERIVED_TWO_PASS", "1")
        else:
            monkeypatch.delenv("SHK_DERIVED_TWO_PASS", raising=False)
        print(f"{shape} {case}: {'two-pass' if two_pass else 'one-pass'} refresh", flush=True)
        hg, hp, hs = hip_state(hip, g_args, raw, p_kw=p_kw, law=lambda g_: law(hip, g_, Cc))
        states = [{n: hs.get(n) for n in FIELD_NAMES}]
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-10, multigrid=True)
        ps = hip.PicardSolver(500, 1e-6, ls, hg)
        sim = hip.Simulation(hg, hs, 3, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(), law(hip, hg, Cc),
                             ps=ps)
        counts = []
        for t in range(3):
            hip.step(sim)
            sim.total_time += sim.dt
            counts.append((ps.converged, ps.last_iter))
            states.append({n: hs.get(n) for n in FIELD_NAMES})
        runs[two_pass] = (states, counts)
        del sim, ps, ls, hs
    (sa, ca), (sb, cb) = runs[True], runs[False]
    assert ca == cb, (ca, cb)
    assert max(last for _, last in ca) > 1  # lean refreshes took part
    for t, (a, b) in enumerate(zip(sa, sb)):
        bad = [n for n in FIELD_NAMES if not np.array_equal(a[n], b[n], equal_nan=True)]
        assert not bad, (t, bad)
