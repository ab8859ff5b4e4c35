"""Pins of the oracle's field sliding laws (CPU): LinearSlidingLaw,
RegularizedCoulombFieldSlidingLaw and RegularizedCoulombV0SlidingLaw
(src/sliding_law.jl:61-90, 270-301, 422-452, 360-396, 509-543, 590-602).

The staggering of C is checked against hand-worked face values; each law
against the regularised-Coulomb law already pinned by its limits (a uniform
field is the scalar law; a field C scales the stress face by face), against
its own limits (V0's Coulomb and power-law ends, its independence from
lambda), and the linear law against its scaling in C and ub."""
import numpy as np
import pytest


def _setup(O, model, ub=1e-6, N0=2e5, nx=40, ny=16, seed=3, A_scale=1.0):
    """A glacier whose GROUNDED cells share one effective pressure N0 and slide
    at a uniform ub in x; ``model(g)`` builds the oracle Model."""
    import shakti_hip.synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=15.0, seed=seed)
    raw = dict(raw)
    raw["ub_x"] = np.full_like(raw["ub_x"], ub)
    raw["A_visc"] = raw["A_visc"] * A_scale
    g = O.Grid(nx, ny, gg.lx, gg.ly)
    p = O.ModelParameters()
    m = model(g)
    s = O.State(g)
    O.set_initial_conditions(s, g, p, m, **raw)
    s.h[...] = s.zb + (s.po - N0) / (p.rho_w * p.g)
    O.update_derived(s, g, p, m)
    return g, p, m, s


def _full_speed_faces(O, s, ub):
    """Interior x-faces between two GROUNDED cells that both slide at ub."""
    gr = (s.mask == O.GROUNDED) & (np.abs(s.abs_ub - ub) <= 1e-15 * ub)
    return gr[1:, :] & gr[:-1, :]


def _c_field(nx, ny, seed=1):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(0.2, 0.9, (nx, ny)))


def test_stagger_hand_values(oracle):
    """3 x 2 cells: outer faces carry their cell's value, inner faces the mean."""
    O = oracle
    g = O.Grid(3, 2, 2.0, 1.0)
    Cc = np.asfortranarray([[1.0, 2.0], [3.0, 4.0], [5.0, 7.0]])  # Cc[ix, iy]
    Cx, Cy = O.stagger_c(g, Cc)
    np.testing.assert_array_equal(Cx, [[1.0, 2.0], [2.0, 3.0], [4.0, 5.5], [5.0, 7.0]])
    np.testing.assert_array_equal(Cy, [[1.0, 1.5, 2.0], [3.0, 3.5, 4.0], [5.0, 6.0, 7.0]])
    # a scalar is a uniform field
    Cx, Cy = O.stagger_c(g, 0.25)
    assert np.all(Cx == 0.25) and np.all(Cy == 0.25) and Cx.shape == (4, 2) and Cy.shape == (3, 3)


@pytest.mark.parametrize("ub", [1e-6, 1e3])
def test_uniform_field_law_is_the_scalar_law(oracle, ub):
    """RegularizedCoulombFieldSlidingLaw(g, c) with a uniform c gives the
    stress of RegularizedCoulombSlidingLaw(c) bit for bit on every face (the
    mean of two equal values is exact)."""
    O = oracle
    _, _, _, a = _setup(O, lambda g: O.Model(sliding_C=0.5), ub=ub)
    _, _, _, b = _setup(O, lambda g: O.Model.field_law(g, "regularized_coulomb_field", 0.5), ub=ub)
    assert np.array_equal(a.taub_x, b.taub_x) and np.array_equal(a.taub_y, b.taub_y)
    assert np.count_nonzero(a.taub_x) > 100


@pytest.mark.parametrize("law", ["regularized_coulomb_field", "regularized_coulomb_v0", "linear"])
def test_field_coefficient_scales_each_face(oracle, law):
    """A non-uniform C enters each face's stress through that face's
    staggered value: taub(C) = Cf * taub(1) for the two regularised-Coulomb
    forms and Cf^2 * taub(1) for the linear law, face by face, on x- and
    y-faces (a transposed or shifted C fails this)."""
    O = oracle
    nx, ny = 40, 16
    Cc = _c_field(nx, ny)
    v0 = 3e-7
    _, _, _, one = _setup(O, lambda g: O.Model.field_law(g, law, 1.0, v0=v0))
    _, _, m, fld = _setup(O, lambda g: O.Model.field_law(g, law, Cc, v0=v0))
    k = 2 if law == "linear" else 1
    for t1, tc, cf in [(one.taub_x, fld.taub_x, m.Cx), (one.taub_y, fld.taub_y, m.Cy)]:
        np.testing.assert_allclose(tc, cf ** k * t1, rtol=4e-16 * 4, atol=0)
    assert np.count_nonzero(fld.taub_x) > 100


def test_linear_law_scaling_and_value(oracle):
    """taub = C^2 N ub: with one N0 on the grounded cells and one ub, the
    full-speed faces carry C^2 N0 ub; doubling ub doubles taub and doubling C
    quadruples it, exactly, on every face; the law has no regularisation, so
    it stays linear at speeds where the regularised-Coulomb law saturates."""
    O = oracle
    N0, c = 2e5, 0.3
    lin = lambda cc: (lambda g: O.Model.field_law(g, "linear", cc))
    _, _, _, a = _setup(O, lin(c), ub=1e-6, N0=N0)
    f = _full_speed_faces(O, a, 1e-6)
    assert f.sum() > 100
    # N on the cells is po - rho_w g (h - zb), N0 to rounding in po
    np.testing.assert_allclose(a.taub_x[1:-1][f], c * c * N0 * 1e-6, rtol=1e-13)
    _, _, _, b = _setup(O, lin(c), ub=2e-6, N0=N0)
    assert np.array_equal(b.taub_x, 2.0 * a.taub_x) and np.array_equal(b.taub_y, 2.0 * a.taub_y)
    _, _, _, d = _setup(O, lin(2 * c), ub=1e-6, N0=N0)
    assert np.array_equal(d.taub_x, 4.0 * a.taub_x)
    _, _, _, e = _setup(O, lin(c), ub=1e3, N0=N0)
    np.testing.assert_allclose(e.taub_x[1:-1][f], c * c * N0 * 1e3, rtol=1e-13)


def test_v0_law_limits(oracle):
    """taub = C N (v / (v + v0))^(1/n) ub / v: v0 << v gives the Coulomb
    limit C N (relative deviation v0 / (n v)); v0 >> v the power law
    C N (v / v0)^(1/n)."""
    O = oracle
    N0, c, ub = 2e5, 0.5, 1e-6
    v0law = lambda v0: (lambda g: O.Model.field_law(g, "regularized_coulomb_v0", c, v0=v0))
    _, _, _, a = _setup(O, v0law(1e-15), ub=ub, N0=N0)
    f = _full_speed_faces(O, a, ub)
    assert f.sum() > 100
    np.testing.assert_allclose(a.taub_x[1:-1][f], c * N0, rtol=1e-9)
    _, _, _, b = _setup(O, v0law(1e3), ub=ub, N0=N0)
    np.testing.assert_allclose(b.taub_x[1:-1][f], c * N0 * (ub / 1e3) ** (1 / 3), rtol=1e-8)


def test_v0_law_ignores_lambda_and_matches_the_field_law_at_its_scale(oracle):
    """The V0 law has no |N|^n lambda term: scaling A (lambda = 1.5 A) leaves
    it unchanged bit for bit while the field law moves; with v0 set to the
    field law's own velocity scale |N0|^3 lambda the two agree on the
    full-speed faces (uniform N0 and lambda there)."""
    O = oracle
    N0, c, ub = 2e5, 0.5, 1e-6
    v0law = lambda v0: (lambda g: O.Model.field_law(g, "regularized_coulomb_v0", c, v0=v0))
    fld = lambda g: O.Model.field_law(g, "regularized_coulomb_field", c)
    _, _, _, a = _setup(O, v0law(2e-6), ub=ub, N0=N0)
    _, _, _, b = _setup(O, v0law(2e-6), ub=ub, N0=N0, A_scale=7.0)
    assert np.array_equal(a.taub_x, b.taub_x) and np.array_equal(a.taub_y, b.taub_y)
    _, _, _, fa = _setup(O, fld, ub=ub, N0=N0)
    _, _, _, fb = _setup(O, fld, ub=ub, N0=N0, A_scale=7.0)
    f = _full_speed_faces(O, fa, ub)
    assert not np.allclose(fa.taub_x[1:-1][f], fb.taub_x[1:-1][f], rtol=1e-3)
    lam_cells = np.unique(getattr(fa, "lambda")[fa.mask == O.GROUNDED])
    assert lam_cells.size == 1
    lam = float(lam_cells[0])
    _, _, _, v = _setup(O, v0law(N0 ** 3 * lam), ub=ub, N0=N0)
    np.testing.assert_allclose(v.taub_x[1:-1][f], fa.taub_x[1:-1][f], rtol=1e-14)


def test_field_law_without_coefficient_is_refused(oracle):
    O = oracle
    with pytest.raises(ValueError):
        O.Model(sliding_law=O.SLIDING_LAWS["linear"]).c()
