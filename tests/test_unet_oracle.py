"""Pins the CPU restatement of the U-Net (oracle/unet_oracle.py) against an
independent implementation (torch CPU autograd, float64) and against central
finite differences, so that the GPU kernels can be checked against it."""
import numpy as np
import pytest

import unet_oracle as U

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _torch_forward(theta, x, slope=0.1):
    """The same network written with torch's conv2d / avg_pool2d / nearest
    interpolate, NCHW, float64."""
    th = torch.as_tensor(theta)
    xs = torch.as_tensor(x).permute(0, 3, 1, 2)
    layers, off = [], 0
    for ca, cb, co, _, _ in U.LAYERS:
        ci = ca + cb
        w = th[off:off + co * 9 * ci].reshape(co, 3, 3, ci).permute(0, 3, 1, 2)  # [co, ci, kh, kw]
        off += co * 9 * ci
        b = th[off:off + co]
        off += co
        layers.append((w, b))
    wh, bh = th[off:off + 16], th[off + 16]

    def conv(l, a, skip=None):
        w, b = layers[l]
        if skip is not None:
            a = torch.cat([F.interpolate(a, scale_factor=2, mode="nearest"), skip], dim=1)
        return F.leaky_relu(F.conv2d(a, w, b, padding=1), slope)

    a0 = conv(0, xs); s0 = conv(1, a0)
    a1 = conv(2, F.avg_pool2d(s0, 2)); s1 = conv(3, a1)
    a2 = conv(4, F.avg_pool2d(s1, 2)); s2 = conv(5, a2)
    a3 = conv(6, F.avg_pool2d(s2, 2)); t3 = conv(7, a3)
    v2 = conv(9, conv(8, t3, s2))
    v1 = conv(11, conv(10, v2, s1))
    v0 = conv(13, conv(12, v1, s0))
    return torch.einsum("bchw,c->bhw", v0, wh) + bh


def _setup(seed=0, B=2, H=32, W=64):
    rng = np.random.default_rng(seed)
    theta = np.zeros(U.num_params())
    off = 0
    for ca, cb, co, _, _ in U.LAYERS:
        ci = ca + cb
        theta[off:off + co * 9 * ci] = rng.normal(0, np.sqrt(2.0 / (9 * ci)), co * 9 * ci)
        off += co * 9 * ci
        theta[off:off + co] = rng.normal(0, 0.05, co)
        off += co
    theta[off:off + 16] = rng.normal(0, 0.25, 16)
    theta[off + 16] = 0.1
    x = rng.normal(0, 1, (B, H, W, 8))
    x[..., 6:] = 0.0
    return theta, x, rng


def test_conv3x3_matches_torch_conv2d():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 5, 7, 8))
    w = rng.normal(size=(16, 9, 8))
    b = rng.normal(size=16)
    ref = F.conv2d(torch.as_tensor(x).permute(0, 3, 1, 2), torch.as_tensor(w).reshape(16, 3, 3, 8).permute(0, 3, 1, 2),
                   torch.as_tensor(b), padding=1).permute(0, 2, 3, 1).numpy()
    np.testing.assert_allclose(U.conv3x3(x, w, b), ref, rtol=1e-12, atol=1e-12)


def test_forward_and_gradient_match_torch_autograd():
    theta, x, rng = _setup()
    out, cache = U.forward(theta, x)
    th = torch.tensor(theta, requires_grad=True)
    ref = _torch_forward(th, x)
    np.testing.assert_allclose(out, ref.detach().numpy(), rtol=1e-10, atol=1e-10)
    dout = rng.normal(size=out.shape)
    (ref * torch.as_tensor(dout)).sum().backward()
    g = U.backward(theta, cache, dout)
    gt = th.grad.numpy()
    assert np.abs(g - gt).max() <= 1e-9 * max(1.0, np.abs(gt).max())
    # every layer receives a nonzero gradient
    off = 0
    for n in U.layer_sizes():
        assert np.abs(g[off:off + n]).max() > 0
        off += n


def test_loss_and_network_gradient_against_finite_differences():
    theta, x, rng = _setup(seed=3)
    t = rng.normal(size=(2, 32, 64))

    def L(th):
        out, _ = U.forward(th, x)
        return U.loss(out, t)[0]

    out, cache = U.forward(theta, x)
    lv, dout = U.loss(out, t)
    g = U.backward(theta, cache, dout)
    # unit directions and a step small enough that no LeakyReLU input
    # changes sign between the two evaluations
    eps = 1e-6
    for k in range(4):
        v = rng.normal(size=theta.shape)
        v /= np.linalg.norm(v)
        fd = (L(theta + eps * v) - L(theta - eps * v)) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-5 * max(1e-4, abs(fd)), (k, fd, g @ v)


def test_loss_is_scale_free_and_zero_for_parallel_output():
    rng = np.random.default_rng(4)
    t = rng.normal(size=(3, 4, 5))
    assert abs(U.loss(2.5 * t, t)[0]) < 1e-14
    assert abs(U.loss(-0.5 * t, t)[0]) < 1e-14
    n = rng.normal(size=t.shape)
    assert abs(U.loss(n, t)[0] - U.loss(7.0 * n, 0.1 * t)[0]) < 1e-12


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 1e-3])
    r = U.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == 1.0
    assert abs(r[4] + 3.140625) < 1e-12
    assert abs(r[5] - 1e-3) <= 2 ** -9 * 1e-3


def test_features_match_a_direct_loop():
    rng = np.random.default_rng(5)
    nx, ny = 7, 5
    aE = rng.uniform(0.1, 2.0, (nx, ny)); aE[-1] = 0
    aN = rng.uniform(0.1, 2.0, (nx, ny)); aN[:, -1] = 0
    cP = rng.uniform(0.0, 0.5, (nx, ny))
    aP = cP.copy()
    aP += aE; aP[1:] += aE[:-1]; aP += aN; aP[:, 1:] += aN[:, :-1]
    r = rng.normal(size=(nx, ny))
    f, s = U.features(aP, aE, aN, cP, r)
    assert f.shape == (32, 64, 8)
    s_ref = np.sqrt(sum(r[i, j] ** 2 / aP[i, j] for i in range(nx) for j in range(ny)) / (nx * ny))
    assert abs(s - s_ref) < 1e-14
    for i in range(nx):
        for j in range(ny):
            d = aP[i, j]
            exp = [r[i, j] / np.sqrt(d) / s,
                   2 * aE[i, j] / np.sqrt(d * aP[i + 1, j]) if i < nx - 1 else 0,
                   2 * aE[i - 1, j] / np.sqrt(d * aP[i - 1, j]) if i > 0 else 0,
                   2 * aN[i, j] / np.sqrt(d * aP[i, j + 1]) if j < ny - 1 else 0,
                   2 * aN[i, j - 1] / np.sqrt(d * aP[i, j - 1]) if j > 0 else 0,
                   cP[i, j] / d, 0, 0]
            np.testing.assert_allclose(f[j, i], exp, rtol=1e-13, atol=1e-15)
    assert not f[ny:].any() and not f[:, nx:].any()


def _consistent_features(rng, B, H, W):
    """Random Jacobi-scaled couplings with the west / south channels equal to
    the neighbours' east / north ones, as the solver writes them."""
    x = np.zeros((B, H, W, 8))
    aE = rng.uniform(0, 1, (B, H, W)); aE[:, :, -1] = 0
    aN = rng.uniform(0, 1, (B, H, W)); aN[:, -1, :] = 0
    x[..., 1], x[..., 3] = aE, aN
    x[..., 2] = U._nb(aE, 0, -1)
    x[..., 4] = U._nb(aN, -1, 0)
    x[..., 5] = rng.uniform(1e-3, 0.1, (B, H, W))
    return x


def test_energy_loss_against_closed_form_and_finite_differences():
    rng = np.random.default_rng(6)
    B, H, W = 2, 6, 9
    x = _consistent_features(rng, B, H, W)
    t = rng.normal(size=(B, H, W))
    x[..., 0] = U.apply_scaled(t, x)  # r = A^ t
    # E(u) = u . A^ u (flux form == operator form)
    assert np.allclose(U.energy(t, x), (t * U.apply_scaled(t, x)).reshape(B, -1).sum(1), rtol=1e-13)
    # zero for any multiple of the solution, positive otherwise
    assert abs(U.loss_energy(3.0 * t, x, t)[0]) < 1e-13
    N = rng.normal(size=(B, H, W))
    lv, g = U.loss_energy(N, x, t)
    assert 0 < lv < 1
    eps = 1e-6
    for k in range(4):
        v = rng.normal(size=N.shape)
        v /= np.linalg.norm(v)
        fd = (U.loss_energy(N + eps * v, x, t)[0] - U.loss_energy(N - eps * v, x, t)[0]) / (2 * eps)
        assert abs(fd - (g * v).sum()) <= 1e-7 * max(1e-3, abs(fd))
