"""The NumPy restatement of the multigrid preconditioner (tests/vcycle_restate.py),
pinned by what the mathematics fixes on a real assembled glacier system (CPU):

* the symmetric red-black V-cycle is a symmetric positive definite operator
  (a swapped smoothing order or a wrong neighbour breaks the symmetry);
* PCG with it converges to 1e-8 in a small fraction of Jacobi-PCG's iterations;
* red-black smoothing needs fewer CG iterations than the same cycle with damped
  Jacobi smoothing (omega 0.95, scale 1.5), the reason it is the default;
* the finest level's coefficients in BF16 cost at most one iteration, the
  reason the library stores them so.

The GPU test (test_gpu_vcycle.py) then checks the library's V-cycle against it."""
import numpy as np
import pytest

from helpers import oracle_state
from vcycle_restate import VCycle, F32, flux_apply

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


def test_vcycle_is_symmetric_positive_definite(system):
    co = system
    M = VCycle(co[0], co[1], co[3])
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
