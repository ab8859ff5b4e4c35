"""Chebyshev bound modes of the oracle, pinned by the dense spectrum (CPU).

The reference's degree-k polynomial (src/preconditioner.jl:239-267) is
positive on (0, lambda_min + lambda_max) and negative just above it, so its
preconditioner is positive definite only if the bounds reach the largest
eigenvalue of D^-1 A.  The Gershgorin lambda_max always does."""
import numpy as np
import pytest


def _glacier(O, nx=32, ny=16, bed_amp=50.0, seed=3):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=bed_amp, seed=seed)
    g = O.Grid(nx, ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = O.Model(kfs=1, sliding_C=0.5)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    co = O.assemble_mfls(s, g, p, 1)
    return g, co


def _dense(co):
    aP, aE, aW, aN, aS = [np.asarray(a).ravel(order="F") for a in co[:5]]
    nx = co[0].shape[0]
    n = aP.size
    A = np.diag(aP)
    for c in range(n):
        ix, iy = c % nx, c // nx
        if ix < nx - 1:
            A[c, c + 1] = -aE[c]
        if ix > 0:
            A[c, c - 1] = -aW[c]
        if c + nx < n:
            A[c, c + nx] = -aN[c]
        if iy > 0:
            A[c, c - nx] = -aS[c]
    return A, aP


def _poly(lam, lmin, lmax, deg):
    """The reference's ldiv! recurrence on a scalar eigenvalue: y = p(lambda)."""
    d, c = (lmax + lmin) / 2, (lmax - lmin) / 2
    r, y = np.ones_like(lam), np.zeros_like(lam)
    a = 1 / d
    p = r.copy()
    y += a * p
    r -= a * lam * p
    ap = a
    for _ in range(2, deg + 1):
        b = (c * ap / 2) ** 2
        a = 1 / (d - b / ap)
        p = r + b * p
        y += a * p
        r -= a * lam * p
        ap = a
    return y


def test_gershgorin_bound_encloses_the_spectrum(oracle):
    O = oracle
    g, co = _glacier(O)
    A, d = _dense(co)
    lam = np.linalg.eigvalsh(A / np.sqrt(np.outer(d, d)))  # D^-1/2 A D^-1/2, similar to D^-1 A
    G = O.gershgorin(co)
    offs = np.abs(A - np.diag(np.diag(A))).sum(axis=1)
    assert G == pytest.approx(1 + np.max(offs / d), rel=1e-15)
    assert lam.min() > 0 and lam.max() <= G
    assert G <= 2.0  # diagonally dominant rows (aP = sum of faces + creep >= sum of couplings)


@pytest.mark.parametrize("start", ["rhs", "smooth", "random"])
def test_d_inner_product_estimate_lies_inside_the_spectrum(oracle, start):
    """The D-inner-product recurrence is a symmetric Lanczos process: its Ritz
    values (the estimate / 0.9 and / 1.1) lie in [lambda_min, lambda_max] of
    D^-1 A, whatever the start vector.  The reference's Euclidean-dot
    recurrence has no such bound (its largest value exceeds the spectrum
    here)."""
    O = oracle
    g, co = _glacier(O)
    A, d = _dense(co)
    nx, ny = co[0].shape
    lam = np.linalg.eigvalsh(A / np.sqrt(np.outer(d, d)))
    X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny), indexing="ij")
    v = {"rhs": co[5], "smooth": np.sin(np.pi * X) * np.sin(np.pi * Y),
         "random": np.random.default_rng(4).standard_normal((nx, ny))}[start]
    lo, hi = O.estimate_eigenvalue_bounds_d(co, np.asfortranarray(v))
    tmin, tmax = lo / 0.9, hi / 1.1
    assert lam.min() * (1 - 1e-10) <= tmin <= tmax <= lam.max() * (1 + 1e-10)
    # 15 steps find the largest eigenvalue well and the smallest within a factor
    assert tmax >= 0.9 * lam.max()
    assert tmin <= 10 * lam.min()
    elo, ehi = O.estimate_eigenvalue_bounds(co, np.asfortranarray(v))
    if start == "smooth":
        assert ehi / 1.1 > lam.max()  # the Euclidean-dot estimate leaves the spectrum


def test_reference_polynomial_positive_exactly_below_lmin_plus_lmax():
    lam = np.linspace(1e-6, 2.5, 250001)
    for lmin, lmax in [(1e-4, 1.9), (0.01, 1.2), (0.3, 0.9)]:
        for deg in (2, 4, 6):
            y = _poly(lam, lmin, lmax, deg)
            below = lam < (lmin + lmax) * (1 - 1e-9)
            above = lam > (lmin + lmax) * (1 + 1e-6)
            assert (y[below] > 0).all()
            assert (y[above] < 0).all()


def test_estimate_that_misses_the_top_makes_the_reference_bounds_indefinite_and_gershgorin_fixes_it(oracle):
    """An estimate whose Krylov space misses the top of the spectrum (here: a
    start vector in the span of the eigenvectors of D^-1 A below 1) returns an
    interval with lambda_min + lambda_max below the largest eigenvalue: with
    the reference's bounds p(D^-1 A) then has negative eigenvalues.  With the
    Gershgorin lambda_max it is positive definite, and the Gershgorin-mode CG
    converges to the dense solution of a full-spectrum right-hand side."""
    O = oracle
    g, co = _glacier(O)
    A, d = _dense(co)
    nx, ny = co[0].shape
    Ssym = A / np.sqrt(np.outer(d, d))
    lam, U = np.linalg.eigh(Ssym)
    low = lam < 1.0
    w = np.random.default_rng(7).uniform(0.5, 1.0, low.sum())
    v = (U[:, low] @ w) / np.sqrt(d)  # right eigenvectors of D^-1 A are D^-1/2 u
    lo, hi = O.estimate_eigenvalue_bounds(co, v.reshape((nx, ny), order="F"))
    assert 0 < lo < hi < 1.1 * 1.0 + 1e-9
    assert lo + hi < lam.max()
    assert _poly(lam, lo, hi, 4).min() < 0  # indefinite preconditioner
    G = O.gershgorin(co)
    assert _poly(lam, lo, G, 4).min() > 0
    assert _poly(np.linspace(1e-9, G, 100001), lo, G, 4).min() > 0
    b = np.random.default_rng(8).standard_normal((nx, ny))
    b = np.asfortranarray(b)
    x_dense = np.linalg.solve(A, b.ravel(order="F"))
    x, st, bb = O.cg_solve(co, b, np.zeros((nx, ny), order="F"), precond="chebyshev", degree=4, atol=0.0,
                           rtol=1e-12, gershgorin=True)
    assert st.solved and not st.zero_curvature and bb[1] == G
    assert bb[0] == O.estimate_eigenvalue_bounds_d(co, b)[0]
    assert np.max(np.abs(x.ravel(order="F") - x_dense)) <= 1e-9 * np.max(np.abs(x_dense))
