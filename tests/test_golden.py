"""The committed config-1 fixture (tests/golden/make_config1_fixture.py):
the oracle must keep reproducing it, and the HIP path must match it to the
plain form's tolerances (h 1e-9, b 1e-8 relative per step, identical Picard
counts, face masks bit-exact)."""
import pathlib
import sys

import numpy as np
import pytest

from helpers import rel_err

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))


@pytest.fixture(scope="module")
def fixture():
    with np.load(GOLDEN / "config1_64x32_48steps.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_oracle_reproduces_config1_fixture(oracle, fixture):
    import make_config1_fixture as M
    out = M.trajectory()
    np.testing.assert_array_equal(out["zb"], fixture["zb"])
    np.testing.assert_array_equal(out["picard"], fixture["picard"])
    for k in ("valid_x", "valid_y"):
        np.testing.assert_array_equal(out[k], fixture[k])
    for t in M.KEEP:
        assert rel_err(out[f"h_{t}"], fixture[f"h_{t}"]) <= 1e-12
        assert rel_err(out[f"b_{t}"], fixture[f"b_{t}"]) <= 1e-12


@pytest.mark.gpu
def test_hip_path_matches_config1_fixture(hip, fixture):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(64, 32, bed_amp=10.0, seed=1)
    np.testing.assert_array_equal(raw["zb"], fixture["zb"])
    g = hip.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    p = hip.ModelParameters()
    s = hip.State(g, p)
    sl = hip.RegularizedCoulombSlidingLaw(0.5)
    hip.set_initial_conditions(s, g, p, sl, **raw)
    np.testing.assert_array_equal(s.get("valid_x"), fixture["valid_x"])
    np.testing.assert_array_equal(s.get("valid_y"), fixture["valid_y"])
    ls = hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13)
    ps = hip.PicardSolver(500, 1e-6, ls, g)
    sim = hip.Simulation(g, s, 48, 3600.0, p, "implicit", [], hip.ConstantMeltInput(), sl, ps=ps)
    for t in range(1, 49):
        hip.step(sim)
        sim.total_time += sim.dt
        assert (int(ps.converged), ps.last_iter) == tuple(fixture["picard"][t - 1]), t
        if f"h_{t}" in fixture:
            assert rel_err(s.get("h"), fixture[f"h_{t}"]) <= 1e-9, t
            assert rel_err(s.get("b"), fixture[f"b_{t}"]) <= 1e-8, t
