"""The library forms x^(1/3.0) as cbrt(x) (1 - d ln x) with d = 1/3 - 1/3.0
= 2^-54 / 3 exactly (pow_third in shk_kernels.hip): the identity and its
constant, restated in numpy, against numpy's own power (CPU; the device path is
held to the oracle by the GPU parity tests)."""
from fractions import Fraction

import numpy as np


def _pow_third(x):
    m, ex = np.frexp(x)
    lnx = (ex.astype(np.float64) + np.log2(m.astype(np.float32)).astype(np.float64)) * 0.69314718055994530942
    y = np.cbrt(x)
    return y + y * (-lnx * (1.0 / 54043195528445952.0))


def test_the_exponent_is_a_third_less_two_to_the_minus_54_over_three():
    assert Fraction(1, 3) - Fraction(1.0 / 3.0) == Fraction(1, 3 * 2**54)
    assert 3 * 2**54 == 54043195528445952


def test_within_two_ulps_of_the_power_over_the_positive_doubles():
    rng = np.random.default_rng(1)
    # log-uniform over every binade below 1 (the sliding stress's range) and a few above
    x = np.exp2(rng.uniform(-1074.0, 8.0, 400_000))
    x = x[x > 0]
    a = np.power(x, 1.0 / 3.0)
    c = _pow_third(x)
    ulps = np.abs(a.view(np.int64) - c.view(np.int64))
    assert ulps.max() <= 2, ulps.max()
    # and the uncorrected cube root is off by up to some 60 ulps for tiny x
    plain = np.abs(a.view(np.int64) - np.cbrt(x).view(np.int64))
    assert plain.max() > 20
