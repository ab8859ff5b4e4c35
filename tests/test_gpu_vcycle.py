"""The library's multigrid preconditioner (one V-cycle, shk_apply_precond)
against the NumPy restatement (tests/vcycle_restate.py) on the same assembled
system, on shapes that between them run every V-cycle kernel: two-column
red-black legs with FP64 and FP32 right-hand sides, the flat legs on odd-width
and small levels, and the LDS tail with the dense coarsest solve.  FP32
inside: the two agree to 5e-4 of |z| (rounding order and the hardware
reciprocal); a wrong neighbour or smoothing order is off by orders of magnitude.
Symmetry r2'M r1 = r1'M r2 is checked on the library's own output."""
import ctypes as C

import numpy as np
import pytest

from helpers import hip_state
from vcycle_restate import VCycle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(130, 66), (520, 130), (1282, 150), (6, 6)])
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
    M = VCycle(aP, aE, aN)  # the library's default smoothing weight and coarse scale
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
        assert err < 5e-4, err
        zs.append(z)
    a, b = float(np.sum(rs[1] * zs[0])), float(np.sum(rs[0] * zs[1]))
    scale = np.sqrt(float(np.sum(rs[0] * zs[0])) * float(np.sum(rs[1] * zs[1])))
    assert abs(a - b) <= 1e-4 * scale, (a, b)
