"""The U-Net kernels (shk_nn.hip) against the CPU restatement
(oracle/unet_oracle.py), tensor by tensor.

Every check recomputes one kernel's output in FP64 from the GPU's own inputs
to that kernel (BF16 values, exact in FP64) and the BF16-rounded weights, so
the only differences left are the GPU's FP32 accumulation order and the BF16
rounding of the output: tolerances are one BF16 step (2^-8 relative) for
BF16 tensors and 1e-4 of the tensor's scale for FP32 sums."""
import numpy as np
import pytest

import unet_oracle as U

pytestmark = pytest.mark.gpu

B, H, W = 2, 32, 128
SLOPE = 0.1
BF16_TOL = 2.0 ** -7


def _close_bf16(gpu, ref, name):
    scale = np.abs(ref).max()
    err = np.abs(gpu - ref) - BF16_TOL * np.abs(ref)
    assert err.max() <= 1e-5 * scale + 1e-30, (name, float(err.max()), float(scale))


def _close_f32(gpu, ref, name, rel=1e-4):
    scale = np.abs(ref).max()
    err = np.abs(gpu - ref).max()
    assert err <= rel * scale + 1e-30, (name, float(err), float(scale))


@pytest.fixture(scope="module")
def run(hip):
    from shakti_hip import _lib as L
    from shakti_hip import learned as LN
    net = LN.UNet(max_batch=B, max_h=H, max_w=W, slope=SLOPE, seed=7, loss="l2")
    rng = np.random.default_rng(11)
    print(f"U-Net {B}x{H}x{W}: forward", flush=True)
    xn = rng.normal(size=(B, H, W, 8))
    x = LN.DeviceArray.from_numpy(xn, "bf16")
    out = net.forward(x).download()
    fw = {n: net.tensor(n, B, H, W) for n, _, _ in L.UNET_TENSORS if not n.startswith("d")}
    tn = (rng.normal(size=(B, H, W)) + 0.5 * out).astype(np.float32)
    t = LN.DeviceArray.from_numpy(tn, "f32")
    print("loss and backward", flush=True)
    loss = net.loss(t, with_grad=True)
    dout = net.tensor("dout", B, H, W)
    net.backward(None)
    bw = {n: net.tensor(n, B, H, W) for n, _, _ in L.UNET_TENSORS if n.startswith("d")}
    theta = net.get_params().astype(np.float64)
    grads = net.get_grads().astype(np.float64)
    return dict(net=net, x=U.bf16_round(xn), out=out, fw=fw, bw=bw, t=tn, loss=loss, dout=dout, theta=theta,
                grads=grads)


def _layer_inputs(fw, l):
    ca, cb, co, upa, lvl = U.LAYERS[l]
    io = [("x",), ("a0",), ("p0",), ("a1",), ("p1",), ("a2",), ("p2",), ("a3",), ("t3", "s2"), ("u2",),
          ("v2", "s1"), ("u1",), ("v1", "s0"), ("u0",)][l]
    X = fw[io[0]].astype(np.float64)
    if len(io) == 2:
        X = np.concatenate([U.up(X), fw[io[1]].astype(np.float64)], axis=-1)
    return X


OUTS = ["a0", "s0", "a1", "s1", "a2", "s2", "a3", "t3", "u2", "v2", "u1", "v1", "u0", "v0"]
DOUTS = ["da0", "ds0", "da1", "ds1", "da2", "ds2", "da3", "dt3", "du2", "dv2", "du1", "dv1", "du0", "dv0"]


def test_forward_every_layer_matches_oracle(run):
    fw, theta = run["fw"], run["theta"]
    Ls = U.unpack(theta)
    for l in range(14):
        w, b = Ls[l]
        X = _layer_inputs(fw, l)
        ref = U.leaky(U.conv3x3(X, U.bf16_round(w), b), SLOPE)
        _close_bf16(fw[OUTS[l]], ref, U.NAMES[l])
    for s_, p_ in (("s0", "p0"), ("s1", "p1"), ("s2", "p2")):
        _close_bf16(fw[p_], U.pool(fw[s_].astype(np.float64)), p_)
    wh, bh = Ls[14]
    _close_f32(run["out"], fw["v0"].astype(np.float64) @ wh + bh, "head", rel=1e-5)
    np.testing.assert_array_equal(run["out"], fw["out"])
    np.testing.assert_array_equal(fw["x"], run["x"])


def test_loss_and_its_gradient_match_oracle(run):
    lv, g = U.loss(run["out"].astype(np.float64), run["t"].astype(np.float64))
    assert abs(run["loss"] - lv) <= 1e-9
    _close_f32(run["dout"], g, "dout", rel=1e-6)


def test_backward_every_tensor_and_gradient_matches_oracle(run):
    fw, bw, theta, grads = run["fw"], run["bw"], run["theta"], run["grads"]
    Ls = U.unpack(theta)
    G = U.unpack(grads)
    f64 = lambda a: a.astype(np.float64)
    # head
    wh, _ = Ls[14]
    dout = f64(run["dout"])
    _close_f32(G[14][0], np.einsum("bhw,bhwc->c", dout, f64(fw["v0"])), "head w grad", rel=1e-5)
    assert abs(G[14][1] - dout.sum()) <= 1e-5 * np.abs(dout).sum()
    _close_bf16(bw["dv0"], U.leaky_grad(fw["v0"], U.bf16_round(dout[..., None] * wh), SLOPE), "dv0")
    # every 3x3 layer: weight / bias gradient from the GPU's own (post-activation) gradient
    for l in range(14):
        X = _layer_inputs(fw, l)
        dz = f64(bw[DOUTS[l]])
        w, _ = Ls[l]
        dx, dw, db = U.conv3x3_bwd(X, U.bf16_round(w), dz)
        _close_f32(G[l][0], dw, U.NAMES[l] + " dW")
        _close_f32(G[l][1], db, U.NAMES[l] + " db")
        # backward-data, checked where the GPU keeps it unmodified or after
        # the consumer's activation derivative
        name = U.NAMES[l]
        if name.startswith("dec") and name.endswith("a"):
            _close_bf16(bw["dcat" + name[3]], dx, name + " dX")
        elif l in (2, 4, 6):
            _close_bf16(bw["dp" + str(l // 2 - 1)], dx, name + " dX")
        elif l in (1, 3, 5, 7):
            _close_bf16(bw["da" + str(l // 2)], U.leaky_grad(fw["a" + str(l // 2)], U.bf16_round(dx), SLOPE),
                        name + " dX")
        elif l in (9, 11, 13):
            u = {9: "u2", 11: "u1", 13: "u0"}[l]
            _close_bf16(bw["d" + u], U.leaky_grad(fw[u], U.bf16_round(dx), SLOPE), name + " dX")
    # concatenation split and pooling backward
    for lvl, (deep, skip, cd) in enumerate([("v1", "s0", 32), ("v2", "s1", 64), ("t3", "s2", 128)]):
        dcat = f64(bw["dcat" + str(lvl)])
        ref_deep = U.up_bwd(dcat[..., :cd])
        _close_bf16(bw["d" + deep], U.leaky_grad(fw[deep], U.bf16_round(ref_deep), SLOPE), "d" + deep)
        ref_skip = U.bf16_round(dcat[..., cd:]) + U.pool_bwd(f64(bw["dp" + str(lvl)]))
        _close_bf16(bw["d" + skip], U.leaky_grad(fw[skip], U.bf16_round(ref_skip), SLOPE), "d" + skip)


def test_energy_loss_and_its_gradient_match_oracle(hip):
    """The energy-norm loss (flux form over the input's couplings) against the
    oracle on the BF16 input the kernels read."""
    from shakti_hip import learned as LN
    net = LN.UNet(max_batch=B, max_h=H, max_w=W, slope=SLOPE, seed=9, loss="energy")
    rng = np.random.default_rng(13)
    x = np.zeros((B, H, W, 8))
    aE = rng.uniform(0, 1, (B, H, W)); aE[:, :, -1] = 0
    aN = rng.uniform(0, 1, (B, H, W)); aN[:, -1, :] = 0
    x[..., 1], x[..., 3] = aE, aN
    x[..., 2], x[..., 4] = U._nb(aE, 0, -1), U._nb(aN, -1, 0)
    x[..., 5] = rng.uniform(1e-3, 0.1, (B, H, W))
    x[..., 0] = rng.normal(size=(B, H, W))
    xd = LN.DeviceArray.from_numpy(x, "bf16")
    out = net.forward(xd).download().astype(np.float64)
    t = rng.normal(size=(B, H, W))
    lv = net.loss(LN.DeviceArray.from_numpy(t, "f32"), with_grad=True)
    ref, g = U.loss_energy(out, U.bf16_round(x), t.astype(np.float32).astype(np.float64))
    assert abs(lv - ref) <= 1e-9
    _close_f32(net.tensor("dout", B, H, W), g, "energy dout", rel=1e-5)


def test_whole_gradient_close_to_fp64_network(run):
    """End to end: the BF16 network against the FP64 oracle network with the
    same FP32 parameters and no rounding anywhere.  Each kernel is exact to
    one BF16 step (tests above); over the ~30 rounded tensors of a forward
    and backward pass that leaves about 1% on the output and about 10% on
    the gradient (measured 10.6%, cosine 0.994)."""
    theta = run["theta"]
    x = run["fw"]["x"].astype(np.float64)
    out, cache = U.forward(theta, x, slope=SLOPE)
    assert np.linalg.norm(out - run["out"]) <= 0.05 * np.linalg.norm(out)
    _, dout = U.loss(out, run["t"].astype(np.float64))
    g = U.backward(theta, cache, dout, slope=SLOPE)
    gg = run["grads"]
    assert np.linalg.norm(g - gg) <= 0.2 * np.linalg.norm(g)
    assert g @ gg >= 0.98 * np.linalg.norm(g) * np.linalg.norm(gg)


def test_adam_step_matches_formula_and_training_is_deterministic(hip):
    from shakti_hip import learned as LN
    rng = np.random.default_rng(5)
    xn = rng.normal(size=(B, H, W, 8))
    x = LN.DeviceArray.from_numpy(xn, "bf16")
    # a target the network can represent: a fixed linear map of its input
    t = LN.DeviceArray.from_numpy(xn[..., 0] + 0.5 * xn[..., 1], "f32")
    runs = []
    for rep in range(2):
        net = LN.UNet(max_batch=B, max_h=H, max_w=W, slope=SLOPE, seed=3)
        th0 = net.get_params().astype(np.float64)
        net.forward(x)
        net.loss(t, with_grad=True)
        net.backward(None)
        g = net.get_grads().astype(np.float64)
        net.adam(1e-3)
        th1 = net.get_params().astype(np.float64)
        m, v = 0.1 * g, 0.001 * g * g
        ref = th0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
        assert np.abs(th1 - ref).max() <= 1e-6 * max(1.0, np.abs(th0).max())
        losses = [net.train_step(x, t, 1e-3) for _ in range(40)]
        runs.append((net.get_params(), losses))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
    # the target is fitted far below the untrained loss
    assert runs[0][1][-1] < 0.5 * runs[0][1][0], runs[0][1][::8]


def test_capacity_and_shape_errors(hip):
    from shakti_hip import learned as LN
    from shakti_hip import _lib as L
    net = LN.UNet(max_batch=1, max_h=32, max_w=64, seed=1)
    x = LN.DeviceArray((1, 32, 96, 8), "bf16")
    with pytest.raises(L.ShaktiHIPError):
        net.forward(x)
    with pytest.raises(L.ShaktiHIPError):
        LN.UNet(max_batch=1, max_h=30, max_w=64)
