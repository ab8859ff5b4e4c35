"""shk_residual_norms_noalloc against shk_residual_norms (runs on an MI355X).

Two calls in a row on one GPU, the second on a smaller grid (the static
scratch keeps the first call's partials): each result equals the context
entry point's bit for bit.  Grids beyond the scratch's capacity are refused
with a message and launch nothing."""
import ctypes as C

import numpy as np
import pytest

from helpers import hip_state

pytestmark = pytest.mark.gpu


def _dev(L, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert L.load().shk_dev_alloc(a.nbytes, C.byref(p)) == 0
    assert L.load().shk_dev_copy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 0) == 0
    return p


def _case(hip, L, nx, ny, seed):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=50.0, seed=seed)
    hg, hp, hs = hip_state(hip, (gg.nx, gg.ny, gg.lx, gg.ly), raw)
    hs._bind(p=hp, kfs=hip.Harmonic())
    L.check(L.load().shk_assemble(hs.ctx, 0, 0.0), hs.ctx)
    sysd = hs.download_system()
    x = np.asfortranarray(np.random.default_rng(seed).standard_normal((nx, ny)) + sysd["rhs"])
    ref = np.zeros(2)
    L.check(L.load().shk_residual_norms(hs.ctx, x.ctypes.data_as(C.POINTER(C.c_double)),
                                        ref.ctypes.data_as(C.POINTER(C.c_double))), hs.ctx)
    flat = lambda a: np.asarray(a).ravel(order="F")
    bufs = [_dev(L, flat(sysd[k])) for k in ("aP", "aE", "aN", "rhs")] + [_dev(L, flat(x)), _dev(L, np.zeros(2))]
    return hs, ref, bufs


def test_noalloc_residual_norms_equal_the_context_entry_point(hip):
    from shakti_hip import _lib as L
    lib = L.load()
    for (nx, ny, seed) in ((1030, 257, 3), (130, 66, 4)):
        hs, ref, bufs = _case(hip, L, nx, ny, seed)
        print(f"noalloc residual norms {nx}x{ny}", flush=True)
        rc = lib.shk_residual_norms_noalloc(nx, ny, *bufs, None)
        assert rc == 0, lib.shk_last_error(None).decode()
        out = np.zeros(2)
        assert lib.shk_dev_copy(out.ctypes.data_as(C.c_void_p), bufs[5], 16, 1) == 0
        assert np.array_equal(out, ref), (out, ref)
        assert ref[0] > 0 and ref[1] > 0
        for p in bufs:
            lib.shk_dev_free(p)


def test_noalloc_refuses_grids_beyond_its_scratch(hip):
    from shakti_hip import _lib as L
    lib = L.load()
    one = _dev(L, np.zeros(2))
    try:
        rc = lib.shk_residual_norms_noalloc(1 << 15, 1 << 15, one, one, one, one, one, one, None)
        assert rc == -1
        msg = lib.shk_last_error(None).decode()
        assert "workgroups" in msg and "32768" in msg, msg
        assert lib.shk_residual_norms_noalloc(1, 5, one, one, one, one, one, one, None) == -1
        assert lib.shk_residual_norms_noalloc(8, 8, None, one, one, one, one, one, None) == -1
    finally:
        lib.shk_dev_free(one)
