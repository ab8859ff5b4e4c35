"""The configurations beyond config 1, on an MI355X.

* 1024 x 256, 200 steps (harmonic K_face, seasonal melt plus 20 moulins,
  implicit gap): the plain FP64 form against the oracle's direct-solve
  trajectory committed in tests/golden (tests/golden/make_config2_fixture.py).
* 4096 x 1024 (bed amplitude 50 m): the precision form (FP32 V-cycle-PCG
  inside FP64 refinement), FP64 CG with the V-cycle and the learned form (the
  trained U-Net's warm start ahead of it) against the plain FP64 form solved
  to 1e-13, where a banded direct solve would need ~34 GB.
* 8192 x 2048 parabolic (e_v 1e-3, bed amplitude 100 m plus a trough): the
  same forms against the plain FP64 form.

Every returned head is also held to its own FP64 residual (the solve log's
true residual), and every plain-FP64 solve must have met its stopping test."""
import pathlib

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
FIELDS = ("h", "b", "N", "q_x", "q_y", "mdot", "K")


def _run(hip, gg, raw, moulin, steps, parabolic=False, **solver):
    from shakti_hip import synthetic as S
    p = hip.ModelParameters(e_v=1e-3) if parabolic else hip.ModelParameters()
    sl = hip.RegularizedCoulombSlidingLaw(0.5)
    s = hip.State(gg, p)
    hip.set_initial_conditions(s, gg, p, sl, **raw)
    ls = hip.CGIterativeSolver(gg, hip.HIPMatrixFreeLinearSystem, atol=0.0, check_every=32, **solver)
    mi = hip.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
    dt = 600.0 if parabolic else 3600.0
    if parabolic:
        sim = hip.Simulation(gg, s, steps, dt, p, "implicit", [], mi, sl, ls=ls, k_face_choice="harmonic")
        ps = None
    else:
        ps = hip.PicardSolver(500, 1e-6, ls, gg)
        sim = hip.Simulation(gg, s, steps, dt, p, "implicit", [], mi, sl, ps=ps, k_face_choice="harmonic")
    s.solve_log()
    log, picard = [], []
    for _ in range(steps):
        hip.step(sim)
        sim.total_time += sim.dt
        log += s.solve_log()
        picard.append((True, 1) if parabolic else (ps.converged, ps.last_iter))
    return s, log, picard


def _healthy(log):
    assert log and all(e["solved"] and not e["zero_curvature"] and not e["tired"] and not e["breakdown"]
                       for e in log)


def test_config2_plain_fp64_matches_the_oracle_trajectory(oracle, hip):
    """1024 x 256, 200 steps: the same Picard (converged, last_iter) every step,
    h within 1e-9 and b within 1e-8 of the oracle at steps 1 and 200, the
    per-step max |h|, sum h, max b and sum b within the same bounds, and the
    face masks bit for bit."""
    from shakti_hip import synthetic as S
    f = np.load(GOLDEN / "config2_1024x256_200steps.npz", allow_pickle=False)
    gg, raw = S.glacier(1024, 256, bed_amp=10.0, seed=2)
    moulin = S.moulin_field(gg, raw["mask"], n=20, seed=2)
    assert np.sum(raw["zb"]) == f["zb_sum"] and np.sum(moulin) == f["moulin_sum"]
    p = hip.ModelParameters()
    sl = hip.RegularizedCoulombSlidingLaw(0.5)
    s = hip.State(gg, p)
    hip.set_initial_conditions(s, gg, p, sl, **raw)
    ls = hip.CGIterativeSolver(gg, hip.HIPMatrixFreeLinearSystem, chebyshev_degree=4, atol=0.0, rtol=1e-13,
                               check_every=32)
    mi = hip.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
    ps = hip.PicardSolver(500, 1e-6, ls, gg)
    sim = hip.Simulation(gg, s, 200, 3600.0, p, "implicit", [], mi, sl, ps=ps, k_face_choice="harmonic")
    s.solve_log()
    worst = {"h": 0.0, "b": 0.0}
    for t in range(1, 201):
        hip.step(sim)
        sim.total_time += sim.dt
        _healthy(s.solve_log())
        assert (int(ps.converged), ps.last_iter) == tuple(f["picard"][t - 1]), t
        h, b = s.get("h"), s.get("b")
        st = f["stats"][t - 1]
        assert abs(np.max(np.abs(h)) - st[0]) <= 1e-9 * st[0], t
        assert abs(np.sum(h) - st[1]) <= 1e-9 * np.abs(h).sum(), t
        assert abs(np.max(b) - st[2]) <= 1e-8 * st[2], t
        assert abs(np.sum(b) - st[3]) <= 1e-8 * st[3], t
        if f"h_{t}" in f.files:
            worst["h"] = max(worst["h"], rel_err(h, f[f"h_{t}"]))
            worst["b"] = max(worst["b"], rel_err(b, f[f"b_{t}"]))
    print(f"config 2: max rel err h {worst['h']:.2e}, b {worst['b']:.2e} at the stored steps", flush=True)
    assert worst["h"] <= 1e-9 and worst["b"] <= 1e-8
    # masks: the oracle's face masks of these inputs
    O = oracle
    g = O.Grid(gg.nx, gg.ny, gg.lx, gg.ly)
    os_ = O.State(g)
    O.set_initial_conditions(os_, g, O.ModelParameters(), O.Model(kfs=1, sliding_C=0.5), **raw)
    for n in ("mask", "valid_x", "valid_y", "ub_x", "ub_y"):
        assert np.array_equal(s.get(n), getattr(os_, n)), n


@pytest.fixture(scope="module")
def c3(hip):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(4096, 1024, bed_amp=50.0, seed=0)
    moulin = S.moulin_field(gg, raw["mask"], n=20, seed=0)
    ref, log, picard = _run(hip, gg, raw, moulin, 2, chebyshev_degree=4, rtol=1e-13)
    _healthy(log)
    out = {f: ref.get(f) for f in FIELDS}
    del ref
    return gg, raw, moulin, out, picard, log


@pytest.mark.parametrize("form", ["fp32_mg", "fp64_mg", "learned_mg"])
def test_config3_forms_against_fp64(hip, c3, form):
    """4096 x 1024, two steps (the first one the initial transient, 16 Picard
    iterations): the same Picard counts as the plain FP64 form and every field
    within the stated bounds of it: h 1e-9, b 1e-8, the derived fields 1e-7.
    learned_mg is the evaluation of the learned form: the committed trained
    U-Net's warm start ahead of FP64 CG with the V-cycle, at the size it is
    evaluated at (it was trained at 512 x 128)."""
    gg, raw, moulin, ref, picard_ref, log_ref = c3
    if form == "fp32_mg":
        kw = dict(multigrid=True, form="fp32_refine", rtol=1e-10, inner_rtol=1e-4)
    elif form == "fp64_mg":
        kw = dict(multigrid=True, form="fp64", rtol=1e-10)
    else:
        from shakti_hip import learned as LN
        H8, W8 = LN.canvas_shape(gg.nx, gg.ny)
        net = LN.UNet.load(LN.WEIGHTS_DIR / "unet_16_32_64_128.npz", 1, H8, W8)
        kw = dict(multigrid=True, form="learned", network=net, rtol=1e-10)
    s, log, picard = _run(hip, gg, raw, moulin, 2, **kw)
    assert picard == picard_ref
    errs = {f: rel_err(s.get(f), ref[f]) for f in FIELDS}
    print(f"{form} vs fp64 at 4096x1024: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert errs["h"] <= 1e-9 and errs["b"] <= 1e-8 and max(errs.values()) <= 1e-7
    assert max(e["true_residual"] for e in log) <= 1e-10
    assert not any(e["zero_curvature"] or e["tired"] or e["breakdown"] for e in log)


def test_config5_parabolic_forms_against_fp64(hip):
    """8192 x 2048 parabolic, three steps of 600 s: the precision form and FP64
    CG with the V-cycle against the plain FP64 form at 1e-13."""
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(8192, 2048, bed_amp=100.0, seed=0, trough_width=2000.0, trough_depth=200.0)
    moulin = S.moulin_field(gg, raw["mask"], n=20, seed=0)
    ref, log, _ = _run(hip, gg, raw, moulin, 3, parabolic=True, chebyshev_degree=4, rtol=1e-13)
    _healthy(log)
    want = {f: ref.get(f) for f in FIELDS}
    del ref
    for form, kw in (("fp32_mg", dict(multigrid=True, form="fp32_refine", rtol=1e-10, inner_rtol=1e-4)),
                     ("fp64_mg", dict(multigrid=True, form="fp64", rtol=1e-10))):
        s, log, _ = _run(hip, gg, raw, moulin, 3, parabolic=True, **kw)
        errs = {f: rel_err(s.get(f), want[f]) for f in FIELDS}
        print(f"{form} vs fp64 at 8192x2048 parabolic: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()),
              flush=True)
        assert errs["h"] <= 1e-9 and errs["b"] <= 1e-8 and max(errs.values()) <= 1e-7
        assert not any(e["zero_curvature"] or e["tired"] or e["breakdown"] for e in log)
        del s
