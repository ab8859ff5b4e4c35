"""The device glacier generator (shk_generate_glacier) against its CPU
restatement (runs on an MI355X): every raw input and the moulin field bit for
bit on grids from 6 x 6 (kernels wider than the grid) to 2048 x 1024; the
initial conditions built from the device inputs against the oracle's from
the restated inputs; two Picard steps with the device moulins against the
oracle's Cholesky steps; and the 8192 x 2048 start timed."""
import time

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

RAW = ["mask", "A_visc", "zb", "zs", "b", "G", "ub_x", "ub_y", "ieb", "taub_x", "taub_y"]

CASES = [
    ((6, 6), dict(seed=1, bed_amp=2.0, corr_len=100e3, n_moulins=4)),  # kernels wrap the grid several times
    ((64, 32), dict(seed=2, bed_amp=10.0, corr_len=2000.0)),
    ((257, 65), dict(seed=3, bed_amp=10.0, corr_len=2000.0, trough_width=2000.0, trough_depth=80.0, n_moulins=20)),
    ((1031, 517), dict(seed=4, bed_amp=50.0, corr_len=700.0, n_moulins=20)),
    ((2048, 1024), dict(seed=5, bed_amp=100.0, corr_len=2000.0, trough_width=2000.0, trough_depth=100.0)),
    ((96, 40), dict(seed=6, bed_amp=5.0, corr_len=0.0, b0=0.004, A=3e-25, ub=2e-6, G=0.06, ieb=2e-8)),
    ((40, 12), dict(seed=9, n_moulins=380)),  # every interior cell: the draw-again rule to the end
]


@pytest.mark.parametrize("shape,kw", CASES, ids=[f"{s[0]}x{s[1]}" for s, _ in CASES])
def test_device_inputs_equal_the_restatement(oracle, hip, shape, kw):
    from shakti_hip import synthetic as S
    nx, ny = shape
    print(f"generator {nx}x{ny} {kw}", flush=True)
    g = hip.Grid(nx, ny, 100e3, 20e3)
    s = hip.State(g)
    S.generate_on_device(s, **kw)
    og = oracle.Grid(nx, ny, 100e3, 20e3)
    raw, moulin = oracle.generate_glacier(og, **kw)
    for n in RAW:
        assert np.array_equal(s.get(n), raw[n]), n
    if moulin is not None:
        from shakti_hip import _lib as L
        # the moulin field read back through ieb mode 2 with a zero uniform part
        L.check(L.load().shk_update_ieb(s.ctx, 2, 0.0), s.ctx, "shk_update_ieb")
        assert np.array_equal(s.get("ieb"), moulin)


def test_initial_conditions_and_steps_from_device_inputs(oracle, hip):
    from shakti_hip import synthetic as S
    nx, ny = 128, 48
    kw = dict(seed=8, bed_amp=10.0, corr_len=2000.0, n_moulins=6)
    g = hip.Grid(nx, ny, 100e3, 20e3)
    p = hip.ModelParameters()
    s = hip.State(g, p)
    S.generate_on_device(s, **kw)
    sl = hip.RegularizedCoulombSlidingLaw(0.5)
    hip.set_initial_conditions_on_device(s, g, p, sl)
    og = oracle.Grid(nx, ny, 100e3, 20e3)
    raw, moulin = oracle.generate_glacier(og, **kw)
    op = oracle.ModelParameters()
    om = oracle.Model(sliding_C=0.5)
    os_ = oracle.State(og)
    oracle.set_initial_conditions(os_, og, op, om, **raw)
    from shakti_hip._lib import FIELD_NAMES
    for n in ["mask", "valid_x", "valid_y", "ub_x", "ub_y"]:
        assert np.array_equal(s.get(n), getattr(os_, n)), n
    bad = {n: rel_err(s.get(n), getattr(os_, n)) for n in FIELD_NAMES}
    bad = {k: v for k, v in bad.items() if v > 1e-13}
    assert not bad, bad
    sv = oracle.Solver(og, "cholesky")
    ls = hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, g)
    mi = hip.SeasonalMoulinMeltInput(None, **S.SEASONAL_C2)  # the moulin field already on the device
    sim = hip.Simulation(g, s, 2, 3600.0, p, "implicit", [], mi, sl, ps=ps)
    for t in range(2):
        u = oracle.seasonal_ieb(sim.total_time, **S.SEASONAL_C2)
        conv, last, _ = oracle.step(sv, os_, og, op, om, 3600.0, ieb_mode=2, ieb_uniform=u, moulin=moulin)
        hip.step(sim)
        sim.total_time += sim.dt
        assert (ps.converged, ps.last_iter) == (conv, last), t
        assert rel_err(s.get("h"), os_.h) <= 1e-9 and rel_err(s.get("b"), os_.b) <= 1e-8, t
        assert rel_err(s.get("ieb"), os_.ieb) <= 1e-15, t


def test_large_start_is_fast(hip):
    """8192 x 2048 (bed amplitude 100 m, 2 km trough, 20 moulins): inputs and
    initial conditions built on the device."""
    from shakti_hip import synthetic as S
    g = hip.Grid(8192, 2048, 100e3, 20e3)
    p = hip.ModelParameters(e_v=1e-3)
    s = hip.State(g, p)
    S.generate_on_device(s, seed=0, bed_amp=100.0, trough_width=2000.0, trough_depth=100.0, n_moulins=20)  # warm
    t0 = time.perf_counter()
    S.generate_on_device(s, seed=0, bed_amp=100.0, trough_width=2000.0, trough_depth=100.0, n_moulins=20)
    t1 = time.perf_counter()
    hip.set_initial_conditions_on_device(s, g, p, hip.RegularizedCoulombSlidingLaw(0.5))
    t2 = time.perf_counter()
    print(f"8192x2048: generator {1e3 * (t1 - t0):.1f} ms, initial conditions {1e3 * (t2 - t1):.1f} ms", flush=True)
    mn, mx, mean = s.stats("zb")
    assert np.isfinite(mn) and np.isfinite(mx) and mx > mn
    assert t1 - t0 < 2.0
