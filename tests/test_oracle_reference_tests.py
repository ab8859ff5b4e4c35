"""The reference's own test invariants, re-proved on the oracle (CPU).

Each test names the reference test it restates; the fixture is the shared
6x6 mixed-mask case (test/linear_solver_test.jl:8-32)."""
import warnings

import numpy as np
import pytest


def _fresh(O, kfs=0, e_v=0.0, b_min=0.0, perturb=True):
    g, p, m, raw = O.test_fixture_6x6(kfs=kfs, e_v=e_v, b_min=b_min)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    if perturb:
        O.perturb_like_reference_tests(s, g, p, m)
    return g, p, m, s


@pytest.mark.parametrize("kfs", [0, 1])
@pytest.mark.parametrize("parabolic", [False, True])
def test_mfls_agrees_with_sals(oracle, kfs, parabolic):
    """test/linear_solver_test.jl:1-83: rhs equal, M exactly symmetric,
    nzval[idxP] == aP and nzval[idxE/W/N/S] == -aE/W/N/S."""
    O = oracle
    g, p, m, s = _fresh(O, kfs=kfs, e_v=1e-3 if parabolic else 0.0)
    aP, aE, aW, aN, aS, rhs = O.assemble_mfls(s, g, p, kfs, parabolic, 60.0)
    M, rhs_s = O.assemble_sals(s, g, p, kfs, parabolic, 60.0)
    np.testing.assert_allclose(rhs_s, rhs.ravel(order="F"), rtol=1e-14, atol=0)
    assert abs(M - M.T).max() == 0.0
    nx, ny = g.nx, g.ny
    Md = M.toarray()
    for j in range(ny):
        for i in range(nx):
            r = i + j * nx
            assert Md[r, r] == pytest.approx(aP[i, j], rel=1e-15)
            if i < nx - 1:
                assert Md[r, r + 1] == pytest.approx(-aE[i, j], rel=1e-15, abs=0)
            if i > 0:
                assert Md[r, r - 1] == pytest.approx(-aW[i, j], rel=1e-15, abs=0)
            if j < ny - 1:
                assert Md[r, r + nx] == pytest.approx(-aN[i, j], rel=1e-15, abs=0)
            if j > 0:
                assert Md[r, r - nx] == pytest.approx(-aS[i, j], rel=1e-15, abs=0)


def test_cg_variants_match_cholesky(oracle):
    """test/linear_solver_test.jl:85-196 for the matrix-free solvers: Jacobi
    and Chebyshev(4) CG with Krylov's default tolerances agree with Cholesky to
    atol 1e-5, and Chebyshev needs no more iterations than Jacobi."""
    O = oracle
    g, p, m, s_lu = _fresh(O)
    O.Solver(g, "cholesky").solve_elliptic(s_lu, p, 0)
    _, _, _, s_j = _fresh(O)
    sj = O.Solver(g, "cg", precond="jacobi")
    sj.solve_elliptic(s_j, p, 0)
    np.testing.assert_allclose(s_j.h, s_lu.h, atol=1e-5, rtol=0)
    _, _, _, s_c = _fresh(O)
    sc = O.Solver(g, "cg", precond="chebyshev", degree=4)
    sc.solve_elliptic(s_c, p, 0)
    np.testing.assert_allclose(s_c.h, s_lu.h, atol=1e-5, rtol=0)
    assert sc.last_stats.niter <= sj.last_stats.niter


def test_solver_constructor_errors(oracle):
    """src/preconditioner.jl:181-182 (degree >= 1, nsteps_estimate >= 2)."""
    O = oracle
    g = O.Grid(6, 6, 1e3, 1e3)
    with pytest.raises(ValueError):
        O.Solver(g, "cg", precond="chebyshev", degree=0)
    with pytest.raises(ValueError):
        O.Solver(g, "cg", precond="chebyshev", nsteps=1)


def test_chebyshev_invalid_estimate_keeps_bounds(oracle):
    """test/linear_solver_test.jl:198-222: a zero rhs makes the estimate NaN;
    the bounds stay at the (1, 1) placeholder and a warning is raised."""
    O = oracle
    n = 36
    # the tridiagonal 4, -1, -1 matrix as a 36x1 stencil grid: aP=4, aE=aW=1
    aP = np.full((n, 1), 4.0, order="F")
    aE = np.full((n, 1), 1.0, order="F")
    aE[-1, 0] = 0.0
    aW = np.full((n, 1), 1.0, order="F")
    aW[0, 0] = 0.0
    z = np.zeros((n, 1), order="F")
    co = (aP, aE, aW, z, z)
    lo, hi, warned = O.update_chebyshev_bounds(co, np.zeros((n, 1), order="F"), bounds=(1.0, 1.0))
    assert warned and lo == 1.0 and hi == 1.0


def test_parabolic_step_moves_h_and_stays_finite(oracle):
    """test/parabolic_solver_test.jl:37-55."""
    O = oracle
    g, p, m, s = _fresh(O, e_v=1e-3, perturb=False)
    h0 = s.h.copy()
    O.parabolic_step(O.Solver(g, "cholesky"), s, g, p, m, 60.0)
    assert not np.allclose(s.h, h0)
    for f in ["h", "N", "Re", "mdot", "K"]:
        assert np.all(np.isfinite(getattr(s, f))), f


def test_parabolic_steady_state_matches_elliptic(oracle):
    """test/parabolic_solver_test.jl:57-87: after 1000 steps of 3600 s the
    parabolic (e_v = 1e-3) and elliptic steady states agree to rtol 1e-4."""
    O = oracle
    g, pe, m, se = _fresh(O, e_v=0.0, b_min=1e-3, perturb=False)
    _, pp, _, sp = _fresh(O, e_v=1e-3, b_min=1e-3, perturb=False)
    sve, svp = O.Solver(g, "cholesky"), O.Solver(g, "cholesky")
    for _ in range(1000):
        O.step(sve, se, g, pe, m, 3600.0)
        O.step(svp, sp, g, pp, m, 3600.0)
    np.testing.assert_allclose(se.h, sp.h, rtol=1e-4)
    np.testing.assert_allclose(se.b, sp.b, rtol=1e-4)


def test_picard_converges_on_fixture(oracle):
    """test/sliding_law_test.jl:173-185 (with the regularised Coulomb law)."""
    O = oracle
    g, p, m, s = _fresh(O, perturb=False)
    conv, last, _ = O.picard_loop(O.Solver(g, "cholesky"), s, g, p, m, 500, 1e-6)
    assert conv and 1 <= last < 500
    assert np.all(np.isfinite(s.h)) and np.all(np.isfinite(s.N))


def test_picard_nonconvergence_is_a_flag(oracle):
    """src/elliptic_solver.jl:173-174: iters exhausted -> converged false,
    last_iter == iters (not an error)."""
    O = oracle
    g, p, m, s = _fresh(O, perturb=False)
    conv, last, _ = O.picard_loop(O.Solver(g, "cholesky"), s, g, p, m, 1, 1e-30)
    assert (not conv) and last == 1
