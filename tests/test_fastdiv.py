"""The V-cycle tail's cell indices: k / w through the FP32 reciprocal of w,
truncate((k + 1/2) * fl(1/w)), in IEEE single precision with round-to-nearest
(the GPU's v_mul_f32 and v_cvt_i32_f32), equals the integer quotient for every
k below 2^16 and every width w up to 4096 (the tail's levels hold at most 4096
cells, its widths at most 4096), and for k below 2^21 over a spread of widths."""
import numpy as np

F32 = np.float32


def _exact(k, widths):
    kf = k.astype(F32) + F32(0.5)
    for w in widths:
        q = (kf * (F32(1.0) / F32(w))).astype(np.int64)
        assert np.array_equal(q, k // w), w


def test_reciprocal_quotient_is_exact():
    # every width for the indices the tail uses (k < 4096) and well beyond
    _exact(np.arange(0, 1 << 16, dtype=np.int64), range(1, 4097))
    # the full range for a spread of widths, powers of two and their neighbours
    widths = sorted(set(range(1, 4097, 61)) | {w + d for w in (2 ** i for i in range(1, 13)) for d in (-1, 0, 1)})
    _exact(np.arange(0, 1 << 21, dtype=np.int64), widths)


def test_the_bound_is_needed():
    """Without the half the quotient fails (k = w m - 0, rounding below m) for
    some width: the half is what keeps every product away from an integer."""
    k = np.arange(0, 1 << 21, dtype=np.int64)
    failures = 0
    for w in (3, 5, 7, 11, 13, 49, 97, 193, 385, 1025, 2049, 3999):
        q = (k.astype(F32) * (F32(1.0) / F32(w))).astype(np.int64)
        failures += int(np.count_nonzero(q != k // w))
    assert failures > 0
