"""CPU restatement (numpy, FP64) of the learned part: the 4-level U-Net, its
backward pass, the scale-free loss and the feature / target maps of the warm
start.  Test infrastructure only: the tests check the GPU kernels of
shakti.jl_amd/csrc/shk_nn.hip against it, and it is itself checked against
torch's CPU autograd and central finite differences (tests/test_unet_oracle.py).

The reference (Shakti.jl) has no learned component; there is no reference
file to cite.  The definitions follow include/shakti_learned.h:
  * images NHWC, (b, y, x, c);
  * 3x3 convolution with zero padding, weights W[cout][tap][cin] with
    tap = 3*(dy+1) + (dx+1), i.e. y[p] = b + sum_tap W[:, tap, :] x[p + (dy, dx)];
  * LeakyReLU(slope) after every 3x3 layer, none after the 1x1 head;
  * 2x2 average pooling, nearest upsampling, decoder input [up(deep), skip].
"""
from __future__ import annotations

import numpy as np

# (cin_a, cin_b, cout, upsampled first input, level) in parameter order
LAYERS = [(8, 0, 16, 0, 0), (16, 0, 16, 0, 0), (16, 0, 32, 0, 1), (32, 0, 32, 0, 1), (32, 0, 64, 0, 2),
          (64, 0, 64, 0, 2), (64, 0, 128, 0, 3), (128, 0, 128, 0, 3), (128, 64, 64, 1, 2), (64, 0, 64, 0, 2),
          (64, 32, 32, 1, 1), (32, 0, 32, 0, 1), (32, 16, 16, 1, 0), (16, 0, 16, 0, 0)]
NAMES = ["enc0a", "enc0b", "enc1a", "enc1b", "enc2a", "enc2b", "enc3a", "enc3b", "dec2a", "dec2b", "dec1a",
         "dec1b", "dec0a", "dec0b", "head"]


def layer_sizes():
    out = []
    for ca, cb, co, _, _ in LAYERS:
        out.append(co * 9 * (ca + cb) + co)
    out.append(17)
    return out


def num_params() -> int:
    return int(sum(layer_sizes()))


def unpack(theta: np.ndarray):
    """Flat parameter vector -> list of (W[co, 9, ci], b[co]) and the head (w[16], b)."""
    theta = np.asarray(theta, dtype=np.float64)
    out, off = [], 0
    for ca, cb, co, _, _ in LAYERS:
        ci = ca + cb
        w = theta[off:off + co * 9 * ci].reshape(co, 9, ci)
        off += co * 9 * ci
        b = theta[off:off + co]
        off += co
        out.append((w, b))
    out.append((theta[off:off + 16], theta[off + 16]))
    return out


def pack(layers) -> np.ndarray:
    parts = []
    for w, b in layers[:-1]:
        parts += [np.asarray(w).ravel(), np.asarray(b).ravel()]
    wh, bh = layers[-1]
    parts += [np.asarray(wh).ravel(), np.atleast_1d(bh)]
    return np.concatenate(parts).astype(np.float64)


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ------------------------------------------------------------------ layers

def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """xs[b, y, x] = x[b, y + dy, x + dx] (zero outside)."""
    B, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv3x3(x: np.ndarray, w: np.ndarray, b=None) -> np.ndarray:
    """y[b, y, x, co] = bias[co] + sum_tap sum_ci w[co, tap, ci] x[b, y+dy, x+dx, ci]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dy, dx) in enumerate(TAPS):
        y += _shift(x, dy, dx) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3x3_bwd(x: np.ndarray, w: np.ndarray, dy: np.ndarray):
    """Gradients (dx, dw, db) of conv3x3 given dy = dL/dy."""
    dx = np.zeros_like(x, dtype=np.float64)
    dw = np.zeros_like(w, dtype=np.float64)
    for t, (sy, sx) in enumerate(TAPS):
        xs = _shift(x, sy, sx)
        dw[:, t, :] = np.einsum("bhwo,bhwi->oi", dy, xs)
        dx += _shift(dy @ w[:, t, :], -sy, -sx)
    return dx, dw, dy.sum(axis=(0, 1, 2))


def leaky(z, s):
    return np.where(z > 0, z, s * z)


def leaky_grad(y, g, s):
    """g * LeakyReLU'(z) evaluated from the output y (same sign as z)."""
    return np.where(y > 0, g, s * g)


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_bwd(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


# ---------------------------------------------------------------- network

def forward(theta, x, slope=0.1, round_bf16=False):
    """Whole forward pass; returns (out[B, H, W], cache of every tensor).
    With round_bf16 the weights and every stored activation are rounded to
    bfloat16 as the GPU stores them (accumulation stays FP64)."""
    L = unpack(theta)
    rnd = bf16_round if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    c = {"x": rnd(x)}

    def conv(l, xa, xb=None):
        w, b = L[l]
        inp = xa if xb is None else np.concatenate([up(xa), xb], axis=-1)
        return rnd(leaky(conv3x3(inp, rnd(w), b), slope))

    c["a0"] = conv(0, c["x"])
    c["s0"] = conv(1, c["a0"])
    c["p0"] = rnd(pool(c["s0"]))
    c["a1"] = conv(2, c["p0"])
    c["s1"] = conv(3, c["a1"])
    c["p1"] = rnd(pool(c["s1"]))
    c["a2"] = conv(4, c["p1"])
    c["s2"] = conv(5, c["a2"])
    c["p2"] = rnd(pool(c["s2"]))
    c["a3"] = conv(6, c["p2"])
    c["t3"] = conv(7, c["a3"])
    c["u2"] = conv(8, c["t3"], c["s2"])
    c["v2"] = conv(9, c["u2"])
    c["u1"] = conv(10, c["v2"], c["s1"])
    c["v1"] = conv(11, c["u1"])
    c["u0"] = conv(12, c["v1"], c["s0"])
    c["v0"] = conv(13, c["u0"])
    wh, bh = L[14]
    c["out"] = c["v0"] @ wh + bh
    return c["out"], c


def backward(theta, c, dout, slope=0.1):
    """Parameter gradient (flat, FP64) of sum(out * dout) for the cache of forward()."""
    L = unpack(theta)
    g = [None] * 15
    wh, bh = L[14]
    g[14] = (np.einsum("bhw,bhwc->c", dout, c["v0"]), dout.sum())
    d = {"v0": dout[..., None] * wh}

    def conv_b(l, out, xa, xb=None):
        w, _ = L[l]
        dz = leaky_grad(c[out], d[out], slope)
        inp = c[xa] if xb is None else np.concatenate([up(c[xa]), c[xb]], axis=-1)
        dx, dw, db = conv3x3_bwd(inp, w, dz)
        g[l] = (dw, db)
        return dx

    d["u0"] = conv_b(13, "v0", "u0")
    dcat = conv_b(12, "u0", "v1", "s0")
    d["v1"], d["s0"] = up_bwd(dcat[..., :32]), dcat[..., 32:]
    d["u1"] = conv_b(11, "v1", "u1")
    dcat = conv_b(10, "u1", "v2", "s1")
    d["v2"], d["s1"] = up_bwd(dcat[..., :64]), dcat[..., 64:]
    d["u2"] = conv_b(9, "v2", "u2")
    dcat = conv_b(8, "u2", "t3", "s2")
    d["t3"], d["s2"] = up_bwd(dcat[..., :128]), dcat[..., 128:]
    d["a3"] = conv_b(7, "t3", "a3")
    d["s2"] = d["s2"] + pool_bwd(conv_b(6, "a3", "p2"))
    d["a2"] = conv_b(5, "s2", "a2")
    d["s1"] = d["s1"] + pool_bwd(conv_b(4, "a2", "p1"))
    d["a1"] = conv_b(3, "s1", "a1")
    d["s0"] = d["s0"] + pool_bwd(conv_b(2, "a1", "p0"))
    d["a0"] = conv_b(1, "s0", "a0")
    conv_b(0, "a0", "x")
    return pack(g)


def loss(out, t):
    """Mean over images of 1 - (N.t)^2 / (|N|^2 |t|^2) and its gradient."""
    B = out.shape[0]
    N = out.reshape(B, -1).astype(np.float64)
    T = t.reshape(B, -1).astype(np.float64)
    c = (N * T).sum(1)
    a = (N * N).sum(1)
    s = (T * T).sum(1)
    Lb = 1.0 - c * c / (a * s)
    grad = (-2.0 * c / (a * s))[:, None] * T + (2.0 * c * c / (a * a * s))[:, None] * N
    return Lb.mean(), (grad / B).reshape(out.shape)


def _nb(u, dy, dx):
    """u at (y+dy, x+dx), zero outside the image (u: [B, H, W])."""
    out = np.zeros_like(u)
    B, H, W = u.shape
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe] = u[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def energy(u, x):
    """E(u) per image: sum c u^2 + sum over east / north faces a (u - u_nb)^2,
    with c = ch5, a_E = ch1 / 2, a_N = ch3 / 2 of the input x [B, H, W, 8]."""
    c, aE, aN = x[..., 5], 0.5 * x[..., 1], 0.5 * x[..., 3]
    dE = u - _nb(u, 0, 1)
    dN = u - _nb(u, 1, 0)
    return (c * u * u + aE * dE * dE + aN * dN * dN).reshape(u.shape[0], -1).sum(1)


def apply_scaled(u, x):
    """A^ u in flux form: c u + sum_faces a_f (u - u_nb), west / south
    couplings from ch2 / ch4."""
    c = x[..., 5]
    out = c * u
    for ch, dy, dx in ((1, 0, 1), (2, 0, -1), (3, 1, 0), (4, -1, 0)):
        out = out + 0.5 * x[..., ch] * (u - _nb(u, dy, dx))
    return out


def loss_energy(out, x, t):
    """Mean over images of 1 - (N.r)^2 / (E(N) E(t)), r = x[..., 0], and its
    gradient (dE/dN = 2 A^ N)."""
    B = out.shape[0]
    N = out.astype(np.float64)
    r = x[..., 0]
    nr = (N * r).reshape(B, -1).sum(1)
    en, et = energy(N, x), energy(t.astype(np.float64), x)
    Lb = 1.0 - nr * nr / (en * et)
    k1 = -2.0 * nr / (en * et)
    k2 = 2.0 * nr * nr / (en * en * et)
    grad = k1[:, None, None] * r + k2[:, None, None] * apply_scaled(N, x)
    return Lb.mean(), grad / B


# ---------------------------------------------------- warm-start features

def canvas_shape(nx: int, ny: int):
    return ((ny + 31) // 32) * 32, ((nx + 63) // 64) * 64


def features(aP, aE, aN, cP, r):
    """Network input for a system in the symmetric face layout (arrays (nx, ny),
    Fortran order as the oracle's State): [H8, W8, 8] float64 before BF16
    rounding, and the scale s."""
    nx, ny = aP.shape
    H8, W8 = canvas_shape(nx, ny)
    s = np.sqrt((r * r / aP).sum() / (nx * ny))
    f = np.zeros((H8, W8, 8))
    sd = np.sqrt(aP)
    rr = r / (sd * s) if s > 0 else np.zeros_like(r)
    E = np.zeros_like(aP)
    Wc = np.zeros_like(aP)
    Nn = np.zeros_like(aP)
    S = np.zeros_like(aP)
    E[:-1] = 2 * aE[:-1] / np.sqrt(aP[:-1] * aP[1:])
    Wc[1:] = 2 * aE[:-1] / np.sqrt(aP[1:] * aP[:-1])
    Nn[:, :-1] = 2 * aN[:, :-1] / np.sqrt(aP[:, :-1] * aP[:, 1:])
    S[:, 1:] = 2 * aN[:, :-1] / np.sqrt(aP[:, 1:] * aP[:, :-1])
    for c, v in enumerate([rr, E, Wc, Nn, S, cP / aP]):
        f[:ny, :nx, c] = v.T
    return f, s
