"""The NumPy restatement of the multigrid preconditioner (tests/vcycle_restate.py),
pinned by what the mathematics fixes on a real assembled glacier system (CPU):

* the symmetric red-black V-cycle is a symmetric positive definite operator
  (a swapped smoothing order or a wrong neighbour breaks the symmetry);
* PCG with it converges to 1e-8 in a small fraction of Jacobi-PCG's iterations;
* red-black smoothing needs fewer CG iterations than the same cycle with damped
  Jacobi smoothing (omega 0.95, scale 1.5), the reason it is the default;
* the finest level's coefficients in BF16 cost at most one iteration, the
  reason the library stores them so;
* the linear transfers: P reproduces constants everywhere and linear fields
  away from the edges, the restriction is its transpose, the lumped coarse
  operator equals the explicit Galerkin product P^T A P with every link laid
  onto its lattice paths, and the cycle with them stays symmetric positive
  definite and needs fewer CG iterations than with aggregation, the reason
  the library uses them below the finest level;
* two sweeps each way on the levels below the finest keep the cycle symmetric
  positive definite and cut CG iterations further.

The GPU test (test_gpu_vcycle.py) then checks the library's V-cycle against it."""
import numpy as np
import pytest

from helpers import oracle_state
from vcycle_restate import VCycle, F32, flux_apply, coarsen_lin, prolong_lin, restrict_lin, library_nu2

NX, NY = 256, 64


@pytest.fixture(scope="module")
def system(oracle):
    import sys, pathlib
    sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "shakti.jl_amd"))
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(NX, NY, bed_amp=50.0, seed=0)
    g, p, m, s = oracle_state(oracle, (gg.nx, gg.ny, gg.lx, gg.ly), raw, kfs=1)
    co = oracle.assemble_mfls(s, g, p, 1)
    return co


def _pcg(oracle, co, M, rtol=1e-8, itmax=3000):
    rng = np.random.default_rng(3)
    b = np.asfortranarray(rng.standard_normal(co[0].shape))
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    g = float(np.sum(r * z))
    r0 = np.sqrt(g)
    p = z.copy()
    for it in range(1, itmax + 1):
        Ap = oracle.matvec(co, p)
        al = g / float(np.sum(p * Ap))
        x += al * p
        r -= al * Ap
        z = M(r)
        gn = float(np.sum(r * z))
        if np.sqrt(gn) <= rtol * r0:
            return it
        p = z + (gn / g) * p
        g = gn
    return itmax


class _JacobiCycle(VCycle):
    """The same hierarchy with one damped-Jacobi sweep down and up."""

    def _cycle(self, b, l):
        c, e, n, D = self.levels[l]
        if l == len(self.levels) - 1:
            return super()._cycle(b, l)
        x = self.omega * b / D
        r = b - flux_apply(c, e, n, x)
        fx, fy = b.shape
        cx, cy = self.levels[l + 1][0].shape
        rp = np.zeros((2 * cx, 2 * cy), F32)
        rp[:fx, :fy] = r
        rc = (rp[0::2, 0::2] + rp[1::2, 0::2]) + (rp[0::2, 1::2] + rp[1::2, 1::2])
        xc = self._cycle(rc.astype(F32), l + 1)
        u = x + self.scale * np.repeat(np.repeat(xc, 2, 0), 2, 1)[:fx, :fy]
        return u + self.omega * (b - flux_apply(c, e, n, u)) / D


@pytest.mark.parametrize("bil_from,nu2,fine", [(None, False, False), (1, False, False), (0, False, False),
                                               (1, True, False), (1, True, True)])
def test_vcycle_is_symmetric_positive_definite(system, bil_from, nu2, fine):
    co = system
    M = VCycle(co[0], co[1], co[3], bil_from=bil_from, nu2=(lambda nx, ny: True) if nu2 else None, nu2_fine=fine)
    rng = np.random.default_rng(0)
    rs = [np.asfortranarray(rng.standard_normal(co[0].shape)) for _ in range(3)]
    zs = [M(r) for r in rs]
    for i in range(3):
        assert float(np.sum(rs[i] * zs[i])) > 0
        for j in range(i):
            a, b = float(np.sum(rs[j] * zs[i])), float(np.sum(rs[i] * zs[j]))
            scale = np.sqrt(float(np.sum(rs[i] * zs[i])) * float(np.sum(rs[j] * zs[j])))
            assert abs(a - b) <= 1e-5 * scale, (i, j, a, b)


def test_red_black_cycle_beats_jacobi_pcg_and_the_jacobi_cycle(oracle, system):
    co = system
    rb = _pcg(oracle, co, VCycle(co[0], co[1], co[3]))
    jc = _pcg(oracle, co, _JacobiCycle(co[0], co[1], co[3], omega=0.95, scale=1.5))
    jacobi = _pcg(oracle, co, lambda r: r / co[0])
    assert rb * 10 < jacobi, (rb, jacobi)
    assert rb < 0.9 * jc, (rb, jc)
    # the finest level's coefficients rounded to BF16 cost at most one iteration
    rb16 = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True))
    assert rb16 <= rb + 1, (rb16, rb)


def _p1d(f):
    """The 1D linear prolongation as a dense f x (f+1)//2 matrix, from its
    definition: fine x takes 3/4 of coarse x // 2 and 1/4 of the coarse
    neighbour on its side, a neighbour outside the grid folding onto x // 2."""
    c = (f + 1) // 2
    P = np.zeros((f, c))
    for x in range(f):
        i0 = x // 2
        i1 = i0 - 1 if x % 2 == 0 else i0 + 1
        if 0 <= i1 < c:
            P[x, i0] += 0.75
            P[x, i1] += 0.25
        else:
            P[x, i0] += 1.0
    return P


SHAPES = [(8, 6), (9, 7), (7, 12), (2, 3), (33, 17), (1, 5)]


@pytest.mark.parametrize("shape", SHAPES)
def test_linear_transfers(shape):
    fx, fy = shape
    cx, cy = (fx + 1) // 2, (fy + 1) // 2
    rng = np.random.default_rng(1)
    # constants everywhere, linear fields between the edge cells
    assert np.allclose(prolong_lin(np.ones((cx, cy)), fx, fy), 1.0, rtol=0, atol=1e-15)
    X = (2 * np.arange(cx) + 1.0)[:, None] + 0 * np.arange(cy)[None, :]  # coarse centres, fine units
    u = prolong_lin(X, fx, fy)
    want = np.arange(fx) + 0.5
    inner = slice(1, fx - 1 if fx % 2 == 0 else fx)
    np.testing.assert_allclose(u[inner, 0], want[inner], rtol=0, atol=1e-13)
    # the matrix of the stated rule, and P^T as the restriction
    xc = rng.standard_normal((cx, cy))
    r = rng.standard_normal((fx, fy))
    Px, Py = _p1d(fx), _p1d(fy)
    np.testing.assert_allclose(prolong_lin(xc, fx, fy), Px @ xc @ Py.T, rtol=0, atol=1e-14)
    np.testing.assert_allclose(restrict_lin(r), Px.T @ r @ Py, rtol=0, atol=1e-14)


def _flux_matrix(c, e, n):
    fx, fy = c.shape
    idx = lambda x, y: x + y * fx
    A = np.diag(c.ravel(order="F").astype(np.float64))
    for y in range(fy):
        for x in range(fx):
            for (x2, y2, a) in ((x + 1, y, e[x, y]), (x, y + 1, n[x, y])):
                if x2 < fx and y2 < fy:
                    i, j = idx(x, y), idx(x2, y2)
                    A[i, i] += a
                    A[j, j] += a
                    A[i, j] -= a
                    A[j, i] -= a
    return A


def _lump(Ac, cx, cy):
    """Every off-diagonal link w = -Ac[i, j] laid onto the coarse faces of the
    x-first and the y-first lattice path between its cells (w / 2 each when
    both paths differ, w on the one path otherwise)."""
    e = np.zeros((cx, cy))
    n = np.zeros((cx, cy))

    def walk(x, y, tx, ty, w, xfirst):
        order = ("x", "y") if xfirst else ("y", "x")
        for ax in order:
            while (x != tx) if ax == "x" else (y != ty):
                if ax == "x":
                    s = 1 if tx > x else -1
                    e[min(x, x + s), y] += w
                    x += s
                else:
                    s = 1 if ty > y else -1
                    n[x, min(y, y + s)] += w
                    y += s

    m = cx * cy
    for i in range(m):
        for j in range(i + 1, m):
            w = -Ac[i, j]
            if w == 0.0:
                continue
            (ix, iy), (jx, jy) = (i % cx, i // cx), (j % cx, j // cx)
            if ix != jx and iy != jy:
                walk(ix, iy, jx, jy, w / 2, True)
                walk(ix, iy, jx, jy, w / 2, False)
            else:
                walk(ix, iy, jx, jy, w, True)
    return e, n


@pytest.mark.parametrize("shape", SHAPES)
def test_lumped_galerkin_coarsening(shape):
    """coarsen_lin against P^T A P formed with dense matrices and lumped link by
    link: c = P^T c (the face part has zero row sums), and the faces' links
    lumped onto the coarse faces."""
    fx, fy = shape
    cx, cy = (fx + 1) // 2, (fy + 1) // 2
    rng = np.random.default_rng(2)
    c = rng.random((fx, fy))
    e = rng.random((fx, fy))
    n = rng.random((fx, fy))
    P = np.kron(_p1d(fy), _p1d(fx))  # cell index x + y fx
    Af = _flux_matrix(np.zeros_like(c), e, n)
    Ac = P.T @ Af @ P
    assert np.allclose(Ac.sum(axis=1), 0.0, atol=1e-12)
    e0, n0 = _lump(Ac, cx, cy)
    c0 = (P.T @ c.ravel(order="F")).reshape((cx, cy), order="F")
    cc, ec, nc = coarsen_lin(c.astype(F32), e.astype(F32), n.astype(F32))
    # the restatement rounds once to FP32 from FP64 sums of the FP32 inputs
    c32, e32, n32 = (a.astype(F32).astype(np.float64) for a in (c, e, n))
    Ac32 = P.T @ _flux_matrix(np.zeros_like(c), e32, n32) @ P
    e1, n1 = _lump(Ac32, cx, cy)
    np.testing.assert_allclose(cc, (P.T @ c32.ravel(order="F")).reshape((cx, cy), order="F"), rtol=1.2e-7, atol=0)
    np.testing.assert_allclose(ec, e1, rtol=1e-6, atol=0)
    np.testing.assert_allclose(nc, n1, rtol=1e-6, atol=0)
    np.testing.assert_allclose(ec, e0, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(cc, c0, rtol=1e-6)
    assert np.all(ec[-1, :] == 0) and np.all(nc[:, -1] == 0)


def test_linear_transfers_cut_cg_iterations(oracle, system):
    co = system
    agg = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True))
    lin1 = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True, bil_from=1))
    lin0 = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True, bil_from=0))
    print(f"PCG iterations: aggregation {agg}, linear below the finest level {lin1}, linear everywhere {lin0}")
    assert lin1 < 0.85 * agg and lin0 < 0.85 * agg, (agg, lin1, lin0)


def test_two_sweeps_cut_cg_iterations(oracle, system):
    """Two red-black sweeps each way on the levels below the finest (the
    library's choice where those levels take the two-row legs), and on the
    finest level too (where it takes the BF16 two-column legs): fewer PCG
    iterations than one sweep, on the same hierarchy, and fewer again."""
    co = system
    one = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True, bil_from=1))
    two = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True, bil_from=1, nu2=lambda nx, ny: True))
    every = _pcg(oracle, co, VCycle(co[0], co[1], co[3], bf16_level0=True, bil_from=1, nu2=lambda nx, ny: True,
                                    nu2_fine=True))
    print(f"PCG iterations: one sweep each way {one}, two below the finest level {two}, two on every level {every}")
    assert every < two < one, (one, two, every)


def test_library_rule_for_two_sweeps():
    """The levels that take two sweeps: more than 4096 cells, even width, and two
    rows per wave at 116 owned columns (ny * ceil(nx / 116) // 4096 <= 2)."""
    assert [library_nu2(nx, ny) for nx, ny in [(2048, 512), (1024, 256), (512, 128), (256, 64), (128, 32)]] == \
        [True, True, True, True, False]
    assert not library_nu2(4096, 1024)      # 36 strips x 1024 rows: eight rows per wave
    assert not library_nu2(513, 256)        # odd width: the flat legs
    assert library_nu2(1026, 512) and library_nu2(2050, 513) and library_nu2(130, 33)
    assert not library_nu2(64, 64)          # 4096 cells: the tail's
