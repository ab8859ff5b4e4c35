"""The glacier generator's CPU restatement (orc_generate_glacier), pinned by
what its definition fixes: SplitMix64's published outputs, the noise
recomputed from the stream in numpy, zero mean and standard deviation
bed_amp to rounding, the Gaussian covariance exp(-r^2 / (2 l^2)) measured on
a large field along both axes, the trough's Gaussian profile, the surface's
closed form, and the moulins' count, placement and range."""
import numpy as np
import pytest

PHI = np.uint64(0x9E3779B97F4A7C15)


def _np_splitmix(z):
    with np.errstate(over="ignore"):
        z = z + PHI
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_stream_is_splitmix64(oracle):
    assert [oracle.splitmix64(k * 0x9E3779B97F4A7C15 % 2 ** 64) for k in range(3)] == \
        [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_unsmoothed_bed_is_the_standardised_noise(oracle):
    """corr_len 0: zb = (w - mean(w)) / sd(w) * bed_amp with w the Irwin-Hall
    sums of the stream, recomputed here from the definition."""
    O = oracle
    g = O.Grid(96, 40, 100e3, 20e3)
    seed = 12345
    raw, _ = O.generate_glacier(g, seed=seed, bed_amp=3.0, corr_len=0.0)
    n = g.nx * g.ny
    c = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) + np.uint64(12) * c * PHI
        w = sum(((_np_splitmix(base + np.uint64(k) * PHI) >> np.uint64(16)).astype(np.float64) * 2.0 ** -48)
                for k in range(12)) - 6.0
    w = w.reshape((g.nx, g.ny), order="F")
    want = (w - w.mean()) / w.std() * 3.0
    np.testing.assert_allclose(raw["zb"], want, rtol=0, atol=1e-13 * 3.0)
    # Irwin-Hall(12): mean 0, variance 1, excess kurtosis -0.1
    big = (w - 0.0).ravel()
    assert abs(big.mean()) < 0.03 and abs(big.var() - 1.0) < 0.05


@pytest.mark.parametrize("amp", [1.0, 50.0])
def test_mean_zero_and_standard_deviation(oracle, amp):
    O = oracle
    g = O.Grid(300, 77, 100e3, 20e3)
    raw, _ = O.generate_glacier(g, seed=7, bed_amp=amp, corr_len=2000.0)
    zb = raw["zb"]
    assert abs(zb.mean()) <= 1e-12 * amp
    assert abs(zb.std() - amp) <= 1e-12 * amp


def test_covariance_is_gaussian_with_the_correlation_length(oracle):
    """Empirical autocorrelation at lags l and 2 l along x and along y (dx and
    dy differ) against exp(-1/2) and exp(-2)."""
    O = oracle
    g = O.Grid(2048, 1024, 100e3, 20e3)  # dx 48.9 m, dy 19.6 m
    ell = 500.0
    raw, _ = O.generate_glacier(g, seed=11, bed_amp=1.0, corr_len=ell)
    zb = raw["zb"]

    def rho(axis, lag_cells):
        a = zb - zb.mean()
        b = np.roll(a, -lag_cells, axis=axis)
        return float(np.mean(a * b) / np.mean(a * a))

    for axis, d in [(0, g.dx), (1, g.dy)]:
        for mult, want in [(1.0, np.exp(-0.5)), (2.0, np.exp(-2.0))]:
            lag = mult * ell / d
            k = int(round(lag))
            # the exact lag in cells is not an integer: compare at the rounded lag
            want_k = np.exp(-0.5 * (k * d / ell) ** 2)
            assert abs(rho(axis, k) - want_k) < 0.04, (axis, mult, rho(axis, k), want_k)


def test_periodic_wrap_beyond_the_grid(oracle):
    """A kernel wider than the grid (6 x 6, s > nx) wraps several times and
    still gives a finite zero-mean field."""
    O = oracle
    g = O.Grid(6, 6, 1e3, 1e3)
    raw, _ = O.generate_glacier(g, seed=1, bed_amp=2.0, corr_len=5000.0)
    assert np.all(np.isfinite(raw["zb"])) and abs(raw["zb"].mean()) < 1e-12
    assert abs(raw["zb"].std() - 2.0) < 1e-12


def test_trough_profile_is_gaussian(oracle):
    """bed_amp 0: zb is the trough alone, constant along x; log(-zb/depth) is
    the parabola -(y - y0)^2 / (2 (w/2)^2) with y0 in [ly/4, 3 ly/4]."""
    O = oracle
    g = O.Grid(64, 400, 100e3, 20e3)
    width, depth = 2000.0, 100.0
    raw, _ = O.generate_glacier(g, seed=5, bed_amp=0.0, trough_width=width, trough_depth=depth)
    zb = raw["zb"]
    assert np.all(zb == zb[:1, :])
    prof = -zb[0, :] / depth
    core = prof > 1e-3
    y = g.y[core]
    c2, c1, c0 = np.polyfit(y, np.log(prof[core]), 2)
    sig = width / 2
    assert c2 == pytest.approx(-1 / (2 * sig ** 2), rel=1e-9)
    y0 = -c1 / (2 * c2)
    assert 0.25 * g.ly <= y0 <= 0.75 * g.ly
    assert prof.max() <= 1.0 and prof.max() > 0.99


def test_surface_and_uniform_fields(oracle):
    O = oracle
    g = O.Grid(50, 20, 100e3, 20e3)
    raw, m = O.generate_glacier(g, seed=2, b0=0.004, A=3e-25, ub=2e-6, G=0.06, ieb=2e-8)
    x = np.arange(g.nx) * g.dx
    np.testing.assert_array_equal(raw["zs"][:, 0], 6.0 * (np.sqrt(x + 5000.0) - np.sqrt(5000.0)) + 1.0)
    assert np.all(raw["zs"] == raw["zs"][:, :1])
    assert np.all(raw["mask"][0, :] == O.OCEAN) and np.all(raw["mask"][1:, :] == O.GROUNDED)
    for name, v in [("b", 0.004), ("A_visc", 3e-25), ("G", 0.06), ("ieb", 2e-8), ("ub_x", 2e-6), ("ub_y", 0.0),
                    ("taub_x", 0.0), ("taub_y", 0.0)]:
        assert np.all(raw[name] == v), name
    assert m is None


def test_moulins_distinct_interior_in_range(oracle):
    O = oracle
    g = O.Grid(40, 12, 100e3, 20e3)
    _, m = O.generate_glacier(g, seed=9, n_moulins=300, qmin=1.0, qmax=10.0)
    nz = np.argwhere(m != 0)
    assert len(nz) == 300  # distinct cells
    assert nz[:, 0].min() >= 1 and nz[:, 0].max() <= g.nx - 2 and nz[:, 1].min() >= 1 and nz[:, 1].max() <= g.ny - 2
    q = m[m != 0] * g.dx * g.dy
    assert q.min() >= 1.0 and q.max() <= 10.0
    # all interior cells: the draw-again rule terminates with every cell taken
    _, full = O.generate_glacier(g, seed=9, n_moulins=(g.nx - 2) * (g.ny - 2))
    assert np.count_nonzero(full[1:-1, 1:-1]) == (g.nx - 2) * (g.ny - 2)
    with pytest.raises(ValueError):
        O.generate_glacier(g, n_moulins=(g.nx - 2) * (g.ny - 2) + 1)


def test_seeds(oracle):
    O = oracle
    g = O.Grid(64, 32, 100e3, 20e3)
    a, ma = O.generate_glacier(g, seed=4, n_moulins=5)
    b, mb = O.generate_glacier(g, seed=4, n_moulins=5)
    c, mc = O.generate_glacier(g, seed=5, n_moulins=5)
    assert np.array_equal(a["zb"], b["zb"]) and np.array_equal(ma, mb)
    assert not np.array_equal(a["zb"], c["zb"]) and not np.array_equal(ma, mc)
