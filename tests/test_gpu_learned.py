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


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
@pytest.mark.parametrize("weights", ["untrained", "trained"])
def test_learned_preconditioner_reaches_the_fp64_solution(oracle, hip, precond, weights):
    """Flexible CG with the network as a preconditioner every 4th iteration
    (composed with the fixed one) stops on the FP64 form's test: the head
    agrees with the FP64 form and the Cholesky oracle to 1e-9."""
    from shakti_hip import learned as LN
    g_args, raw = _glacier(257, 65, seed=8, amp=30.0)
    H8, W8 = LN.canvas_shape(257, 65)
    if weights == "trained":
        if not LN.DEFAULT_WEIGHTS.exists():
            pytest.skip("no trained weights committed")
        net = LN.UNet.load(LN.DEFAULT_WEIGHTS, 1, H8, W8)
    else:
        net = LN.UNet(max_batch=1, max_h=H8, max_w=W8, seed=3)
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    res = {}
    for form, every in (("fp64", 0), ("learned", 4)):
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, form=form,
                                   network=net if form == "learned" else None, learned_precond_every=every, **kw)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        res[form] = (hs.get("h"), ls.ws.stats)
    st = res["learned"][1]
    print(f"{precond} {weights}: CG iterations fp64 {res['fp64'][1].niter}, learned preconditioner {st.niter}, "
          f"fallback {st.fallback}", flush=True)
    assert st.solved
    assert rel_err(res["learned"][0], res["fp64"][0]) <= 1e-9
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    co = oracle.assemble_mfls(os_, g, p, 1)
    assert rel_err(res["learned"][0], oracle.cholesky_solve(co, co[5])) <= 1e-9


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
def test_learned_preconditioner_with_a_zero_network_is_plain_pcg(hip, precond):
    """A network whose output is zero gives e = 0, a rejected step and z = P(r):
    flexible CG then takes the same steps as PCG (its beta equals Krylov's for
    a fixed preconditioner, up to rounding), so the head agrees with the FP64
    form and the iteration count differs by at most the convergence test a
    learned iteration skips (and one more for rounding)."""
    from shakti_hip import learned as LN
    g_args, raw = _glacier(200, 48, seed=12, amp=20.0)
    H8, W8 = LN.canvas_shape(200, 48)
    zero = LN.UNet(max_batch=1, max_h=H8, max_w=W8, seed=None)
    zero.set_params(np.zeros(zero.n_params, dtype=np.float32))
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=3)
    res = {}
    for form, every in (("fp64", 0), ("learned", 6)):
        hg, hp, hs = hip_state(hip, g_args, raw)
        ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-10, form=form,
                                   network=zero if form == "learned" else None, learned_precond_every=every, **kw)
        hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
        res[form] = (hs.get("h"), ls.ws.stats)
    a, b = res["fp64"][1], res["learned"][1]
    print(f"{precond}: iterations {a.niter} / {b.niter}", flush=True)
    assert b.solved and b.fallback & 1  # the zero warm start is rejected too
    assert abs(b.niter - a.niter) <= 2
    assert rel_err(res["learned"][0], res["fp64"][0]) <= 1e-9


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
def test_learned_preconditioner_stall_falls_back_to_pcg(oracle, hip, precond):
    """A stall threshold of 0 counts the first learned batch as stalled: the
    solve goes on as plain PCG, reports it (fallback bit 1) and still meets
    the FP64 form's test."""
    from shakti_hip import learned as LN
    g_args, raw = _glacier(130, 40, seed=5, amp=25.0)
    H8, W8 = LN.canvas_shape(130, 40)
    net = LN.UNet(max_batch=1, max_h=H8, max_w=W8, seed=4)
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    hg, hp, hs = hip_state(hip, g_args, raw)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, form="learned", network=net,
                               learned_precond_every=2, stall_ratio=0.0, **kw)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
    st = ls.ws.stats
    assert st.solved and st.fallback & 2
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    co = oracle.assemble_mfls(os_, g, p, 1)
    assert rel_err(hs.get("h"), oracle.cholesky_solve(co, co[5])) <= 1e-9
    # without the forced stall the flag stays clear
    hg, hp, hs = hip_state(hip, g_args, raw)
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-11, form="learned", network=net,
                               learned_precond_every=2, stall_ratio=1e3, **kw)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
    assert ls.ws.stats.solved and not ls.ws.stats.fallback & 2


@pytest.mark.parametrize("precond", ["multigrid", "chebyshev"])
def test_recorded_cg_pairs_hold_residual_and_remaining_error(oracle, hip, precond):
    """Each recorded CG pair (input: the iteration's scaled residual r_k;
    target: its scaled remaining error e_k = x_final - x_k) satisfies
    A e_k = r_k: sqrt(aP) channel 0 and A (t / sqrt(aP)) agree up to the
    BF16 rounding of the input (the scale s_k cancels)."""
    from shakti_hip import learned as LN
    nx, ny = 130, 40
    g_args, raw = _glacier(nx, ny, seed=6, amp=25.0)
    H8, W8 = LN.canvas_shape(nx, ny)
    zero = LN.UNet(1, H8, W8, seed=None)
    zero.set_params(np.zeros(zero.n_params, dtype=np.float32))
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    hg, hp, hs = hip_state(hip, g_args, raw)
    cap = 40
    x = LN.DeviceArray((cap, H8, W8, 8), "bf16")
    t = LN.DeviceArray((cap, H8, W8), "f32")
    rec = LN.PairRecorder(hs, x, t, 0, cap, kind="cg")
    ls = hip.CGIterativeSolver(hg, hip.HIPMatrixFreeLinearSystem, atol=0.0, rtol=1e-13, form="learned",
                               network=zero, learned_precond_every=2, stall_ratio=1e3, **kw)
    hip.solve_elliptic_linear_system(ls, hs, hg, hp, hip.Harmonic())
    n = rec.stop()
    niter = ls.ws.stats.niter
    # the first batch is plain; then one learned iteration per batch of 2
    assert 1 <= n <= min(cap, niter // 2), (n, niter)
    g, p, m, os_ = oracle_state(oracle, g_args, raw, kfs=1)
    aP, aE, aW, aN, aS, rhs = oracle.assemble_mfls(os_, g, p, 1)
    xh, th = x.download(), t.download()
    # pairs of the first half of the solve: their remaining error stands well
    # above the rounding of x_final - x_k (near convergence it does not)
    for k in range(0, max(1, n // 2), max(1, n // 8)):
        f0 = xh[k, :ny, :nx, 0].T.astype(np.float64)
        e = th[k, :ny, :nx].T.astype(np.float64) / np.sqrt(aP)
        Ae = aP * e
        Ae[:-1] -= aE[:-1] * e[1:]; Ae[1:] -= aW[1:] * e[:-1]
        Ae[:, :-1] -= aN[:, :-1] * e[:, 1:]; Ae[:, 1:] -= aS[:, 1:] * e[:, :-1]
        r = np.sqrt(aP) * f0
        err = np.linalg.norm(Ae - r) / np.linalg.norm(r)
        print(f"{precond} pair {k}/{n}: |A e - r| / |r| = {err:.2e}", flush=True)
        assert err <= 2e-2, (k, err)
    # the target channel is zero outside the grid
    assert (th[:n, ny:, :] == 0).all() and (th[:n, :, nx:] == 0).all()


def test_learned_preconditioner_arguments(hip):
    from shakti_hip import _lib as L
    from shakti_hip import learned as LN
    g = hip.Grid(64, 32, 1e5, 2e4)
    net = LN.UNet(1, 32, 64, seed=0)
    with pytest.raises(L.ShaktiHIPError):
        hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, learned_precond_every=4)  # fp64 form
    with pytest.raises(L.ShaktiHIPError):
        hip.CGIterativeSolver(g, hip.HIPMatrixFreeLinearSystem, form="learned", network=net,
                              learned_precond_every=4)  # Jacobi: nothing to compose with
    p = hip.ModelParameters()
    s = hip.State(g, p)
    lib = L.load()
    for every, ratio in ((1, 1.0), (-2, 1.0), (5000, 1.0), (4, -1.0), (4, float("nan"))):
        assert lib.shk_set_learned_precond(s.ctx, every, ratio) != 0
    assert lib.shk_set_learned_precond(s.ctx, 0, 1.0) == 0


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
