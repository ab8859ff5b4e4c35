"""The learned warm start inside the solver: recorded training pairs against
the oracle, the learned form's solutions against the FP64 form and the
Cholesky oracle, and what the trained network gains."""
import numpy as np
import pytest

from helpers import hip_state, oracle_state, rel_err

import unet_oracle as U

pytestmark = pytest.mark.gpu


def _glacier(nx, ny, seed, amp=10.0):
    from shakti_hip import synthetic as S
    gg, raw = S.glacier(nx, ny, bed_amp=amp, seed=seed)
    return (gg.nx, gg.ny, gg.lx, gg.ly), raw


@pytest.mark.parametrize("shape", [(64, 32), (100, 40)])
def test_recorded_pair_matches_oracle_features_and_target(oracle, hip, shape):
    from shakti_hip import learned as LN
    g_args, raw = _glacier(*shape, seed=4)
    hg, hp, hs = hip_state(hip, g_args, raw)
    H8, W8 = LN.canvas_shape(*shape)
    x = LN.DeviceArray.from_numpy(np.full((2, H8, W8, 8), 7.0), "bf16")
    t = LN.DeviceArray.from_numpy(np.full((2, H8, W8), 7.0), "f32")
    rec = LN.PairRecorder(hs, x, t, 0, 1)
    h0 = hs.get("h")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-12)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())  # capacity 1: not recorded
    assert rec.stop() == 1
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    aP, aE, aW, aN, aS, rhs = oracle.assemble_mfls(os_, g, p, 1)
    h_ref = oracle.cholesky_solve((aP, aE, aW, aN, aS, rhs), rhs)
    A_h = aP * h0
    A_h[:-1] -= aE[:-1] * h0[1:]; A_h[1:] -= aW[1:] * h0[:-1]
    A_h[:, :-1] -= aN[:, :-1] * h0[:, 1:]; A_h[:, 1:] -= aS[:, 1:] * h0[:, :-1]
    r = rhs - A_h
    f_ref, s = U.features(aP, aE, aN, aP - (aE + aW + aN + aS), r)
    xh, th = x.download(), t.download()
    f = xh[0]
    err = np.abs(f - f_ref) - 2.0 ** -8 * np.abs(f_ref)
    assert err.max() <= 1e-6 * max(1.0, np.abs(f_ref).max()), float(err.max())
    t_ref = np.zeros((H8, W8))
    t_ref[:shape[1], :shape[0]] = (np.sqrt(aP) * (h_ref - h0) / s).T
    assert rel_err(th[0], t_ref) <= 1e-6
    # the second slot was never written
    assert (xh[1] == 7.0).all() and (th[1] == 7.0).all()


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
def test_learned_form_reaches_the_fp64_solution(oracle, hip, precond):
    """With any network (here an untrained one) the learned form stops on the
    same residual test as the FP64 form, measured from h: the head agrees with
    the FP64 form and the Cholesky oracle."""
    from shakti_hip import learned as LN
    g_args, raw = _glacier(257, 65, seed=8, amp=30.0)
    H8, W8 = LN.canvas_shape(257, 65)
    net = LN.UNet(max_batch=1, max_h=H8, max_w=W8, seed=1)
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    res = {}
    for form in ("fp64", "learned"):
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, form=form,
                                   network=net if form == "learned" else None, **kw)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        res[form] = (hs.get("h"), ls.ws.stats)
    st = res["learned"][1]
    assert st.solved
    # the reference norm is the one at h, as for the FP64 form
    assert abs(st.rnorm0 - res["fp64"][1].rnorm0) <= 1e-12 * res["fp64"][1].rnorm0
    assert rel_err(res["learned"][0], res["fp64"][0]) <= 1e-9
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    co = oracle.assemble_mfls(os_, g, p, 1)
    assert rel_err(res["learned"][0], oracle.cholesky_solve(co, co[5])) <= 1e-9


def test_learned_form_needs_a_network(hip):
    from shakti_hip import _lib as L
    g = hip.Grid(64, 32, 1e5, 2e4)
    with pytest.raises(L.ShaktiHIPError):
        hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, form="learned")


def test_learned_form_picard_trajectory_matches_oracle(oracle, hip):
    """Config-1 style run (64x32, Cholesky oracle) with the learned warm start:
    Picard converges in the same number of iterations and h, b stay within
    the plain form's tolerances."""
    from shakti_hip import learned as LN
    g_args, raw = _glacier(64, 32, seed=1)
    g, p, m, os_ = oracle_state(oracle, g_args, raw)
    hg, hp, hs = hip_state(hip, g_args, raw)
    sv = oracle.Solver(g, "cholesky")
    net = LN.UNet.load(LN.DEFAULT_WEIGHTS, 1, 32, 64) if LN.DEFAULT_WEIGHTS.exists() else LN.UNet(1, 32, 64, seed=2)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-12,
                               form="learned", network=net)
    ps = hip.PicardSolver(500, 1e-6, ls, hg)
    sim = hip.Simulation(hg, hs, 12, 3600.0, hp, "implicit", [], hip.ConstantMeltInput(),
                         hip.RegularizedCoulombSlidingLaw(0.5), ps=ps)
    for t in range(12):
        oracle.step(sv, os_, g, p, m, 3600.0)
        hip.step(sim)
        assert ps.converged
        assert rel_err(hs.get("h"), os_.h) <= 1e-9, t
        assert rel_err(hs.get("b"), os_.b) <= 1e-8, t


def test_trained_network_gain_on_held_out_glaciers(hip):
    """What the committed, trained network gains over the method without it,
    on 512x128 glaciers it never saw (seeds outside the training set):
    * prediction: on pairs recorded from the plain solver on two glaciers
      of the training family it never saw, the trained network's
      scale-free L2 loss is far below the untrained network's (validation
      loss after training: 0.14, untrained: 0.64; on the large corrections
      of the first timestep after initialisation it is 0.34 against 0.61);
    * solve: FP64 CG with the multigrid preconditioner needs no more CG
      iterations from the learned warm start than from h (measured about
      1.5% fewer; the energy-minimising step never makes the start worse)."""
    from shakti_hip import learned as LN
    from shakti_hip import synthetic as S
    from shakti_hip import train_unet as T
    if not LN.DEFAULT_WEIGHTS.exists():
        pytest.skip("no trained weights committed")
    nx, ny = 512, 128
    H8, W8 = LN.canvas_shape(nx, ny)
    x = LN.DeviceArray((24, H8, W8, 8), "bf16")
    t = LN.DeviceArray((24, H8, W8), "f32")
    idx = []
    for k, variant in enumerate((64, 65)):  # training used variants 0..63
        g, raw, moulin = LN.glacier_variant(variant, nx, ny)
        p = hip.ModelParameters()
        sl = hip.RegularizedCoulombSlidingLaw(0.5)
        s = hip.State(g, p)
        hip.set_initial_conditions(s, g, p, sl, **raw)
        ls = hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, multigrid=True, atol=0.0, rtol=1e-10)
        ps = hip.PicardSolver(500, 1e-6, ls, g)
        sim = hip.Simulation(g, s, 10, 3600.0, p, "implicit", [], hip.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2),
                             sl, ps=ps, k_face_choice="harmonic")
        hip.step(sim)  # the pairs start after the first timestep, as most training pairs do
        sim.total_time += sim.dt
        rec = LN.PairRecorder(s, x, t, 12 * k, 12)
        while rec.count < 12:
            hip.step(sim)
            sim.total_time += sim.dt
        idx += list(range(12 * k, 12 * k + rec.stop()))
    trained = LN.UNet.load(LN.DEFAULT_WEIGHTS, 8, H8, W8)
    untrained = LN.UNet(8, H8, W8, seed=0)
    l_tr = LN.evaluate(trained, x, t, idx, 8)
    l_un = LN.evaluate(untrained, x, t, idx, 8)
    print(f"held-out loss: trained {l_tr:.3f}, untrained {l_un:.3f}", flush=True)
    assert l_tr < 0.3 and l_tr < 0.5 * l_un
    rows = T.gain(LN.DEFAULT_WEIGHTS, nx, ny, [7003], 2, "multigrid")
    print(rows, flush=True)
    assert rows[0]["cg_learned"] <= rows[0]["cg_plain"]
    assert rows[0]["h_rel_diff"] <= 1e-9
