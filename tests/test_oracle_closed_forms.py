"""Pins of the oracle by what the mathematics fixes (CPU).

Closed forms of the per-cell kernels, dense linear algebra (numpy/scipy) for
the operator, Cholesky and Chebyshev, LAPACK's tridiagonal eigensolver for the
Sturm bisection, and brute-force mask rules on several mask layouts.  Each
check is chosen so that a plausible restatement slip (a dropped term, a wrong
sign or index, a transposed operand) fails it."""
import numpy as np
import pytest
import scipy.linalg as sla


def _state(O, raw_fn=None, nx=48, ny=20, seed=0, kfs=0, perturb=True):
    import shakti_hip.synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=15.0, seed=seed)
    g = O.Grid(nx, ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = O.Model(kfs=kfs)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    if perturb:
        rng = np.random.default_rng(seed)
        s.h += rng.uniform(-5.0, 5.0, s.h.shape)
        O.update_derived(s, g, p, m)
    return g, p, m, s


def test_flux_quadratic_closed_form(oracle):
    """Re solves omega Re^2 + Re - D = 0 with D = b^3 g |grad h| / (12 nu^2),
    |q| = nu Re and q flows down the head gradient (src/water_flux.jl:192-216)."""
    O = oracle
    g, p, m, s = _state(O)
    for b, dh, q, Re in [(s.b_x, s.dhdx, s.q_x, s.Re_x), (s.b_y, s.dhdy, s.q_y, s.Re_y)]:
        D = b ** 3 * p.g * np.abs(dh) / (12 * p.nu ** 2)
        np.testing.assert_allclose(p.omega * Re ** 2 + Re, D, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(np.abs(q), p.nu * Re, rtol=1e-15)
        nz = (dh != 0) & (Re > 0)
        assert np.all(np.sign(q[nz]) == -np.sign(dh[nz]))


def test_pressure_and_effective_pressure(oracle):
    O = oracle
    g, p, m, s = _state(O)
    np.testing.assert_allclose(s.pw, p.rho_w * p.g * (s.h - s.zb), rtol=1e-15)
    grounded = s.mask == O.GROUNDED
    np.testing.assert_allclose(s.N[grounded], (s.po - s.pw)[grounded], rtol=1e-15)
    assert np.all(s.N[~grounded] == 0)


def test_gradients_masked_and_boundary_untouched(oracle):
    O = oracle
    g, p, m, s = _state(O)
    dx = np.diff(s.h, axis=0) / g.dx * s.valid_x[1:-1, :]
    np.testing.assert_allclose(s.dhdx[1:-1, :], dx, rtol=1e-15, atol=0)
    assert np.all(s.dhdx[0, :] == 0) and np.all(s.dhdx[-1, :] == 0)
    dy = np.diff(s.pw, axis=1) / g.dy * s.valid_y[:, 1:-1]
    np.testing.assert_allclose(s.dpwdy[:, 1:-1], dy, rtol=1e-14, atol=1e-9)


def test_transmissivity_and_cell_reynolds(oracle):
    O = oracle
    g, p, m, s = _state(O)
    np.testing.assert_allclose(s.K * (12 * p.nu * (1 + p.omega * s.Re)), s.b ** 3 * p.g, rtol=1e-14)
    Re = (s.Re_x[:-1, :] + s.Re_x[1:, :] + s.Re_y[:, :-1] + s.Re_y[:, 1:]) / 4
    np.testing.assert_allclose(s.Re, Re, rtol=1e-15)


def _uniform_state(O, ub=1e-6, N0=2e5, flat_head=False, G_scale=1.0, nx=40, ny=16, seed=3):
    """A glacier whose GROUNDED cells share one effective pressure N0 (h set to
    zb + (po - N0) / (rho_w g)) or, with flat_head, one head (no gradient, so
    no water flow), with a uniform basal speed ub in x on the faces the mask
    keeps; A (so lambda = 1.5 A) is uniform."""
    import shakti_hip.synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=15.0, seed=seed)
    raw = dict(raw)
    raw["ub_x"] = np.full_like(raw["ub_x"], ub)
    raw["G"] = raw["G"] * G_scale
    g = O.Grid(nx, ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = O.Model(kfs=0, sliding_C=0.5)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    if flat_head:
        s.h[...] = np.max(s.zb) + 100.0
    else:
        s.h[...] = s.zb + (s.po - N0) / (p.rho_w * p.g)
    O.update_derived(s, g, p, m)
    return g, p, m, s


def _interior_grounded_x_faces(O, s, ub):
    """x-faces 1..nx-1 between two GROUNDED cells that both slide at the full
    speed ub (cells next to a face the mask zeroed have a lower |ub|, and the
    face speed v is the average of the two cells' |ub|)."""
    gr = (s.mask == O.GROUNDED) & (np.abs(s.abs_ub - ub) <= 1e-15 * ub)
    return gr[1:, :] & gr[:-1, :]


def test_regularised_coulomb_limits(oracle):
    """The regularised-Coulomb law taub = C N (v / (v + |N|^n lambda))^(1/n) ub / v
    (src/sliding_law.jl:213-256): with one N on every grounded cell its face
    value is that N whatever the averaging, so the law itself is pinned by its
    limits on the faces whose face speed v equals ub.  Fast sliding
    (v >> |N|^n lambda): the Coulomb limit C N.  Slow sliding: the power law
    C (v / lambda)^(1/n), independent of N.  In between |taub| <= C |N| there
    (elsewhere the bound is C |N| ub_f / v, with v the average of the two
    cells' speeds)."""
    O = oracle
    N0 = 2e5
    lam = 1.5 * 2.5e-25
    C = 0.5
    _, _, _, fast = _uniform_state(O, ub=1e3, N0=N0)
    f = _interior_grounded_x_faces(O, fast, 1e3)
    assert f.sum() > 100
    t = fast.taub_x[1:-1][f]
    # |N|^3 lambda / v = 3e-9: the relative deviation from C N is a third of that
    np.testing.assert_allclose(t, C * N0, rtol=2e-9)
    # v / (|N|^3 lambda) = 3.3e-8: the power law to a relative 1.1e-8
    _, _, _, slow = _uniform_state(O, ub=1e-16, N0=N0)
    t = slow.taub_x[1:-1][f]
    np.testing.assert_allclose(t, C * (1e-16 / lam) ** (1 / 3), rtol=5e-8)
    # the slow limit does not depend on N
    _, _, _, slow2 = _uniform_state(O, ub=1e-16, N0=2 * N0)
    np.testing.assert_allclose(slow2.taub_x[1:-1][f], t, rtol=5e-8)
    for ub in (1e-9, 1e-6, 1e-3):
        _, _, _, s = _uniform_state(O, ub=ub, N0=N0)
        assert np.all(np.abs(s.taub_x[1:-1][f]) <= C * N0 * (1 + 1e-15))
        assert np.all(np.abs(s.taub_x[1:-1][f]) > 0)
        # where a cell's speed is lowered the stress can exceed C N, by ub_f / v at most
        v = (s.abs_ub[1:] + s.abs_ub[:-1]) / 2
        pos = v > 0
        bound = C * N0 * np.abs(s.ub_x[1:-1][pos]) / v[pos]
        assert np.all(np.abs(s.taub_x[1:-1][pos]) <= bound * (1 + 1e-14))
    # no speed, no stress
    _, _, _, still = _uniform_state(O, ub=0.0, N0=N0)
    assert np.all(still.taub_x == 0) and np.all(still.taub_y == 0)


def test_transmissivity_laminar_limit_and_melt_of_geothermal_heat_alone(oracle):
    """With no head gradient there is no flow (q = 0, Re = 0), so K is the
    laminar cubic law b^3 g / (12 nu) (src/water_flux.jl:221-234); with no
    sliding either, every heat term but the geothermal flux vanishes and
    mdot = G / L (src/melt_rate.jl:106-122), linear in G."""
    O = oracle
    _, p, _, s = _uniform_state(O, ub=0.0, flat_head=True)
    assert np.all(s.q_x == 0) and np.all(s.q_y == 0) and np.all(s.Re == 0)
    np.testing.assert_allclose(s.K, s.b ** 3 * p.g / (12 * p.nu), rtol=1e-15, atol=0)
    assert np.all(s.shear == 0) and np.all(s.potential == 0) and np.all(s.sensible == 0)
    np.testing.assert_allclose(s.mdot, s.G / p.L, rtol=1e-15, atol=0)
    _, _, _, s3 = _uniform_state(O, ub=0.0, flat_head=True, G_scale=3.0)
    np.testing.assert_allclose(s3.mdot, 3 * s.mdot, rtol=1e-15)
    # sliding adds frictional heat |ub . taub| only: mdot rises above G/L where ub != 0
    _, _, _, sl = _uniform_state(O, ub=1e-6, flat_head=True)
    assert np.all(sl.potential == 0) and np.all(sl.sensible == 0)
    np.testing.assert_allclose(p.L * sl.mdot - sl.G, sl.shear, rtol=1e-12, atol=1e-15 * np.max(sl.G))
    assert np.all(sl.mdot >= s.mdot) and np.any(sl.mdot > s.mdot)


def test_explicit_gap_step_is_forward_euler(oracle):
    """The explicit scheme (src/gap_height.jl:127-135) is a forward-Euler step
    b + dt f(b): linear in dt (b(2 dt) - b = 2 (b(dt) - b) to rounding where no
    clamp acts), and consistent with the implicit backward-Euler step, which
    it matches to first order: |b_E - b_I| = O(dt^2)."""
    O = oracle
    g, p, m, s0 = _state(O)

    def stepped(scheme, dt):
        s = s0.copy()
        mm = O.Model(kfs=m.kfs, gap_scheme=scheme, sliding_C=m.sliding_C)
        O.step_b(s, g, p, mm, dt)
        return s.b

    gr = s0.mask == O.GROUNDED
    b0 = s0.b
    e1, e2 = stepped(1, 60.0), stepped(1, 120.0)
    keep = gr & (e1 > 0) & (e2 > 0)
    assert keep.sum() > 100
    np.testing.assert_allclose((e2 - b0)[keep], 2 * (e1 - b0)[keep], rtol=1e-9, atol=1e-18)
    assert np.all(e1[~gr] == b0[~gr])
    d = []
    for dt in (400.0, 200.0, 100.0):
        d.append(np.max(np.abs(stepped(1, dt) - stepped(0, dt))[keep]))
    assert 3.5 < d[0] / d[1] < 4.5 and 3.5 < d[1] / d[2] < 4.5, d


def test_gap_height_backward_euler(oracle):
    """Implicit b satisfies b_new (1 + dt A |N|^2 N) = b + dt (mdot/rho_i + beta |ub|)
    on GROUNDED cells before clamping; other cells untouched (src/gap_height.jl:117-125)."""
    O = oracle
    g, p, m, s = _state(O)
    b0, N, A, mdot, beta, ub = s.b.copy(), s.N.copy(), s.A_visc.copy(), s.mdot.copy(), s.beta.copy(), s.abs_ub.copy()
    dt = 3600.0
    O.step_b(s, g, p, m, dt)
    gr = (s.mask == O.GROUNDED) & (s.b > 0)
    lhs = s.b * (1 + dt * A * N ** 2 * N)
    rhs = b0 + dt * (mdot / p.rho_i + beta * ub)
    np.testing.assert_allclose(lhs[gr], rhs[gr], rtol=1e-13)
    assert np.all(s.b[s.mask != O.GROUNDED] == b0[s.mask != O.GROUNDED])
    np.testing.assert_allclose(s.beta, np.maximum(0, (p.br - s.b) / p.lr), rtol=1e-15)
    np.testing.assert_allclose(s.b_x[1:-1], (s.b[1:] + s.b[:-1]) / 2, rtol=1e-15)
    assert np.all(s.b_x[0] == s.b[0]) and np.all(s.b_x[-1] == s.b[-1])


def _masks():
    """Mask layouts: the reference test fixture, and the three MaskVariants
    shapes of docs/src/examples/MaskVariants.jl:24-66 (simple, barrier,
    semicircle) on a small odd grid."""
    out = []
    nx = ny = 15
    m = np.zeros((nx, ny))
    m[-1, :] = 1.0
    m[0, :] = 2.0
    m[:, 0] = 3.0
    out.append(m)
    simple = np.zeros((nx, ny))
    simple[0, :] = 1.0
    out.append(simple)
    barrier = simple.copy()
    barrier[nx // 2, 2:ny - 2] = 3.0
    out.append(barrier)
    xx, yy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    semi = np.where((xx - 0) ** 2 + (yy - ny // 2) ** 2 > (0.8 * nx) ** 2, 3.0, 0.0)
    semi[0, :] = 1.0
    out.append(semi)
    return out


@pytest.mark.parametrize("k", range(4))
def test_mask_bookkeeping_brute_force(oracle, k):
    """valid faces and ub masking against a direct loop over faces
    (src/mask.jl:62-150): integer-exact."""
    O = oracle
    mask = _masks()[k]
    nx, ny = mask.shape
    g, p, m, raw = O.test_fixture_6x6()
    g = O.Grid(nx, ny, 1e3, 1e3)
    raw = dict(mask=mask, A_visc=np.full((nx, ny), 5e-25), zb=np.zeros((nx, ny)), zs=np.full((nx, ny), 300.0),
               b=np.full((nx, ny), 0.01), G=np.full((nx, ny), 0.05), ub_x=np.full((nx + 1, ny), 1e-6),
               ub_y=np.full((nx, ny + 1), 2e-7), ieb=np.zeros((nx, ny)), taub_x=np.zeros((nx + 1, ny)),
               taub_y=np.zeros((nx, ny + 1)))
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    for ix in range(nx + 1):
        for iy in range(ny):
            if ix in (0, nx):
                vx, ux = 1.0, 0.0
            else:
                a, b = mask[ix - 1, iy], mask[ix, iy]
                vx = 0.0 if (a == 3.0 or b == 3.0) else 1.0
                ux = 1e-6 if (a == 0.0 and b == 0.0) else 0.0
            assert s.valid_x[ix, iy] == vx and s.ub_x[ix, iy] == ux
    for ix in range(nx):
        for iy in range(ny + 1):
            if iy in (0, ny):
                vy, uy = 1.0, 0.0
            else:
                a, b = mask[ix, iy - 1], mask[ix, iy]
                vy = 0.0 if (a == 3.0 or b == 3.0) else 1.0
                uy = 2e-7 if (a == 0.0 and b == 0.0) else 0.0
            assert s.valid_y[ix, iy] == vy and s.ub_y[ix, iy] == uy
    assert np.all(s.b[mask != 0.0] == 0.0)


def test_assembly_row_structure(oracle):
    """Identity rows on OCEAN/LAND/OTHER_BASIN cells; a GROUNDED row's diagonal
    minus its face conductances is the creep reaction n rho_w g A |N|^2 b; the
    Dirichlet head enters the rhs folded through the face conductance."""
    O = oracle
    g, p, m, raw = O.test_fixture_6x6()
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    O.perturb_like_reference_tests(s, g, p, m)
    aP, aE, aW, aN, aS, rhs = O.assemble_mfls(s, g, p, 0)
    for i in range(g.nx):
        for j in range(g.ny):
            mk = s.mask[i, j]
            if mk != O.GROUNDED:
                assert aP[i, j] == 1 and aE[i, j] == aW[i, j] == aN[i, j] == aS[i, j] == 0
                if mk == O.OTHER_BASIN:
                    assert rhs[i, j] == s.h[i, j]
                continue
            faces = []
            for di, dj, d2 in [(1, 0, g.dx2), (-1, 0, g.dx2), (0, 1, g.dy2), (0, -1, g.dy2)]:
                ii, jj = i + di, j + dj
                if not (0 <= ii < g.nx and 0 <= jj < g.ny):
                    continue
                mo = s.mask[ii, jj]
                kf = 0.0 if mo == O.OTHER_BASIN else (s.K[i, j] if mo in (O.OCEAN, O.LAND)
                                                      else (s.K[i, j] + s.K[ii, jj]) / 2)
                faces.append(kf / d2)
            react = p.n * p.rho_w * p.g * s.A_visc[i, j] * s.N[i, j] ** 2 * s.b[i, j]
            assert aP[i, j] == pytest.approx(sum(faces) + react, rel=1e-14)


def test_operator_equals_dense_matrix(oracle):
    O = oracle
    g, p, m, s = _state(O, nx=40, ny=17, kfs=1)
    co = O.assemble_mfls(s, g, p, 1)
    M, rhs = O.assemble_sals(s, g, p, 1)
    x = np.random.default_rng(1).standard_normal((g.nx, g.ny))
    y = O.matvec(co, x)
    xf = x.ravel(order="F")
    scale = abs(M) @ np.abs(xf)  # per-row sum of |terms|: bounds the rounding of the row's sum
    assert np.all(np.abs(y.ravel(order="F") - M @ xf) <= 1e-14 * scale)


def test_cholesky_matches_lapack(oracle):
    """Band Cholesky vs numpy's dense solve (6x6) and LAPACK's banded SPD
    solver scipy.linalg.solveh_banded (64x32, both orderings)."""
    O = oracle
    g, p, m, raw = O.test_fixture_6x6()
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    co = O.assemble_mfls(s, g, p, 0)
    M, rhs = O.assemble_sals(s, g, p, 0)
    x = O.cholesky_solve(co, co[5])
    np.testing.assert_allclose(x.ravel(order="F"), np.linalg.solve(M.toarray(), rhs), rtol=1e-12)
    for nx, ny in [(64, 32), (24, 50)]:
        g, p, m, s = _state(O, nx=nx, ny=ny)
        co = O.assemble_mfls(s, g, p, 0)
        M, rhs = O.assemble_sals(s, g, p, 0)
        n, bw = nx * ny, nx
        ab = np.zeros((bw + 1, n))
        Md = M.toarray()
        for k in range(bw + 1):
            ab[bw - k, k:] = np.diagonal(Md, k)
        xr = sla.solveh_banded(ab, rhs)
        x = O.cholesky_solve(co, co[5])
        np.testing.assert_allclose(x.ravel(order="F"), xr, rtol=1e-11, atol=1e-9 * np.abs(xr).max())


def test_tridiagonal_extremes_match_lapack(oracle):
    rng = np.random.default_rng(7)
    for m_ in [2, 5, 15, 40]:
        a = rng.uniform(0.1, 3.0, m_)
        e = rng.uniform(-1.0, 1.0, m_ - 1)
        ev = sla.eigvalsh_tridiagonal(a, e)
        lo, hi = oracle.tridiag_extreme_eigs(a, e)
        assert lo == pytest.approx(ev[0], rel=1e-13, abs=1e-14)
        assert hi == pytest.approx(ev[-1], rel=1e-13, abs=1e-14)
    lo, hi = oracle.tridiag_extreme_eigs(np.array([1.0, np.nan]), np.array([0.5]))
    assert np.isnan(lo) and np.isnan(hi)


def _laplacian(nx, ny):
    aP = np.full((nx, ny), 4.0, order="F")
    aE = np.ones((nx, ny), order="F")
    aE[-1, :] = 0
    aW = np.ones((nx, ny), order="F")
    aW[0, :] = 0
    aN = np.ones((nx, ny), order="F")
    aN[:, -1] = 0
    aS = np.ones((nx, ny), order="F")
    aS[:, 0] = 0
    return (aP, aE, aW, aN, aS)


def test_chebyshev_placeholder_is_jacobi_richardson(oracle):
    """With the placeholder bounds (1, 1), degree-k ldiv! equals
    sum_{j<k} (I - D^-1 A)^j D^-1 x (src/preconditioner.jl:239-267)."""
    O = oracle
    g, p, m, s = _state(O, nx=20, ny=9)
    co = O.assemble_mfls(s, g, p, 0)
    M, _ = O.assemble_sals(s, g, p, 0)
    Md = M.toarray()
    D = np.diag(Md)
    x = np.random.default_rng(2).standard_normal(g.nx * g.ny)
    for k in [1, 2, 4]:
        y = O.cheb_apply(co, x.reshape((g.nx, g.ny), order="F"), k, 1.0, 1.0).ravel(order="F")
        B = np.eye(len(D)) - Md / D[:, None]
        t = x / D
        acc = np.zeros_like(t)
        for _ in range(k):
            acc += t
            t = B @ t
        np.testing.assert_allclose(y, acc, rtol=1e-12, atol=1e-12 * np.abs(acc).max())


def _scalar_cheb(k, d, c, lam, first_beta_factor=1.0):
    """1 - lam p(lam) for the ldiv! recurrence applied to one eigencomponent."""
    a = 1 / d
    s, p = 1.0, 1.0
    y = a * p
    s = s - a * lam * p
    ap = a
    for j in range(2, k + 1):
        b = (c * ap / 2) ** 2 * (first_beta_factor if j == 2 else 1.0)
        a = 1 / (d - b / ap)
        p = s + b * p
        y += a * p
        s = s - a * lam * p
        ap = a
    return 1 - lam * y


def test_chebyshev_eigenvector_response(oracle):
    """On an eigenvector v of D^-1 A (lam known in closed form for the 5-point
    Laplacian) the preconditioner returns p(lam) v; degree 1 gives p = 1/d.
    The reference recurrence (src/preconditioner.jl:241-264) takes its first
    beta as (c alpha_0 / 2)^2; with that beta doubled -- Saad's Chebyshev
    semi-iteration -- the residual polynomial is exactly
    T_k((d-lam)/c) / T_k(d/c), which pins the rest of the recurrence."""
    O = oracle
    nx, ny = 17, 9
    co = _laplacian(nx, ny)
    i, j = 3, 2
    xs = np.sin(np.pi * (i + 1) * (np.arange(nx) + 1) / (nx + 1))
    ys = np.sin(np.pi * (j + 1) * (np.arange(ny) + 1) / (ny + 1))
    v = np.asfortranarray(np.outer(xs, ys))
    lam = 1 - (np.cos(np.pi * (i + 1) / (nx + 1)) + np.cos(np.pi * (j + 1) / (ny + 1))) / 2
    lmin, lmax = 0.05, 1.9
    d, c = (lmax + lmin) / 2, (lmax - lmin) / 2
    for k in [1, 2, 3, 4, 6]:
        y = O.cheb_apply(co, 4.0 * v, k, lmin, lmax)   # x = D v so D^-1 x = v
        pl = y[2, 3] / v[2, 3]
        np.testing.assert_allclose(y, pl * v, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(1 - lam * pl, _scalar_cheb(k, d, c, lam), rtol=1e-10, atol=1e-13)
        T = np.polynomial.chebyshev.Chebyshev.basis(k)
        np.testing.assert_allclose(_scalar_cheb(k, d, c, lam, 2.0), T((d - lam) / c) / T(d / c), rtol=1e-10,
                                   atol=1e-13)
        if k == 1:
            assert pl == pytest.approx(1 / d, rel=1e-15)


def test_lanczos_bounds_bracket_spectrum(oracle):
    """For symmetric D^-1 A (constant diagonal) the CG-Lanczos Ritz values lie
    inside the spectrum, so the 0.9/1.1-scaled bounds enclose it; and the
    estimate equals an independent numpy CG + LAPACK tridiagonal eigensolve."""
    O = oracle
    nx, ny = 12, 10
    co = _laplacian(nx, ny)
    rhs = np.asfortranarray(np.random.default_rng(3).standard_normal((nx, ny)))
    lo, hi = O.estimate_eigenvalue_bounds(co, rhs, 15)
    lam = [1 - (np.cos(np.pi * a / (nx + 1)) + np.cos(np.pi * b / (ny + 1))) / 2
           for a in range(1, nx + 1) for b in range(1, ny + 1)]
    assert lo / 0.9 >= min(lam) - 1e-12 and hi / 1.1 <= max(lam) + 1e-12
    assert lo <= min(lam) * 1.0 + 0.5 and hi >= max(lam) * 0.9
    # independent restatement with numpy + scipy's LAPACK tridiagonal solver
    g, p, m, s = _state(O, nx=30, ny=12)
    co = O.assemble_mfls(s, g, p, 0)
    M, _ = O.assemble_sals(s, g, p, 0)
    Md = M.toarray()
    D = np.diag(Md)
    b = co[5].ravel(order="F")
    r, pp = b.copy(), b.copy()
    gam = r @ r
    al, be = [], []
    for k in range(15):
        Ap = (Md @ pp) / D
        al.append(gam / (pp @ Ap))
        if k < 14:
            r = r - al[-1] * Ap
            gn = r @ r
            be.append(gn / gam)
            pp = r + be[-1] * pp
            gam = gn
    dd = [1 / al[0]] + [1 / al[k] + be[k - 1] / al[k - 1] for k in range(1, 15)]
    ee = [np.sqrt(be[k]) / al[k] for k in range(14)]
    ev = sla.eigvalsh_tridiagonal(np.array(dd), np.array(ee))
    lo, hi = O.estimate_eigenvalue_bounds(co, co[5], 15)
    assert lo == pytest.approx(0.9 * ev[0], rel=1e-8)
    assert hi == pytest.approx(1.1 * ev[-1], rel=1e-8)


@pytest.mark.parametrize("precond", ["jacobi", "chebyshev"])
def test_cg_tight_residual_and_warm_start(oracle, precond):
    """Tight CG reaches ||b - A x|| / ||b|| <= 1e-12 from the warm start; a
    warm start that is already the solution returns with 0 iterations."""
    O = oracle
    g, p, m, s = _state(O, nx=64, ny=32, kfs=1)
    co = O.assemble_mfls(s, g, p, 1)
    b = co[5]
    x, st, _ = O.cg_solve(co, b, s.h, precond=precond, atol=0.0, rtol=1e-14)
    res = b - O.matvec(co, x)
    assert np.linalg.norm(res) / np.linalg.norm(b) <= 1e-12
    assert st.solved
    x_exact = O.cholesky_solve(co, b)
    x2, st2, _ = O.cg_solve(co, b, x_exact, precond=precond)
    assert st2.niter == 0


def test_seasonal_forcing_closed_form(oracle):
    """src/melt_input.jl:86-94 at the window edges and outside it, and the
    1-day sinusoid used by the 1024x256 workload."""
    O = oracle
    spy = 365 * 86400.0
    assert O.seasonal_ieb(0.4 * spy) == pytest.approx((493.75 - 492.75) / spy, rel=1e-12)
    assert O.seasonal_ieb(0.55 * spy) == pytest.approx((493.75 + 492.75) / spy, rel=1e-12)
    assert O.seasonal_ieb(0.1 * spy) == pytest.approx(1.0 / spy, rel=1e-15)
    from shakti_hip.synthetic import SEASONAL_C2, seasonal_c2
    for t in [0.0, 3600.0, 43200.0, 90000.0]:
        v = O.seasonal_ieb(t, **{k: SEASONAL_C2[k] for k in ["t_start", "t_end", "amplitude", "offset", "i_min",
                                                           "seconds_per_year"]})
        assert v == pytest.approx(1e-8 + 1e-7 * 0.5 * (1 - np.cos(2 * np.pi * t / 86400)), rel=1e-12)
        assert seasonal_c2(t) == pytest.approx(v, rel=1e-12)
