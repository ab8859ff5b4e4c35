"""The learned warm start (include/shakti_learned.h): a 4-level U-Net
(16/32/64/128 channels, 3x3 convolutions, BF16 weights and activations, FP32
accumulation) whose forward, backward and Adam steps are the library's own
gfx950 kernels.

* ``UNet``: the network handle (parameters, forward, loss, backward, Adam,
  train step, internal tensors for checks).
* ``PairRecorder``: makes every elliptic solve of a State write one training
  pair (network input of the Picard iteration, scaled correction) into device
  buffers.
* ``DeviceArray``: device buffers from the library's own allocator (a second
  HIP runtime in the same process, such as the one bundled with torch,
  cannot share the GPU with the library's).
* ``generate_pairs`` / ``train``: the solver-generated data set and the
  training loop.

The reference has no learned component; the solver hook sits behind the same
seam as the rest (solve_elliptic_linear_system!, src/linear_solver.jl:915-925).
"""
from __future__ import annotations

import ctypes as C
import math
import pathlib
import time
import weakref

import numpy as np

from . import _lib as L

WEIGHTS_DIR = pathlib.Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "unet_16_32_64_128.npz"


def canvas_shape(nx: int, ny: int):
    """(H8, W8): the grid padded to the network's image shape (ny up to a
    multiple of 32, nx up to a multiple of 64)."""
    return ((ny + 31) // 32) * 32, ((nx + 63) // 64) * 64


def _finalize_net(handle):
    if handle.value:
        L.load().shk_unet_destroy(handle)
        handle.value = None


def _finalize_buf(handle):
    if handle.value:
        L.load().shk_dev_free(handle)
        handle.value = None


def to_bf16_bits(a) -> np.ndarray:
    """float -> bfloat16 bit patterns (uint16), rounding to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class DeviceArray:
    """A device buffer (shk_dev_alloc) with a shape and an element type,
    "bf16" (host side: uint16 bit patterns) or "f32"."""

    def __init__(self, shape, dtype: str = "f32"):
        if dtype not in ("bf16", "f32"):
            raise ValueError(dtype)
        self.shape = tuple(int(v) for v in shape)
        self.dtype = dtype
        self.itemsize = 2 if dtype == "bf16" else 4
        self.nbytes = int(np.prod(self.shape)) * self.itemsize
        self._h = C.c_void_p()
        L.check(L.load().shk_dev_alloc(self.nbytes, C.byref(self._h)), None, "shk_dev_alloc")
        weakref.finalize(self, _finalize_buf, self._h)

    @property
    def ptr(self) -> C.c_void_p:
        return self._h

    def row_bytes(self) -> int:
        return self.nbytes // self.shape[0]

    def at(self, k: int) -> C.c_void_p:
        """Address of entry k along the first axis."""
        return C.c_void_p(self._h.value + k * self.row_bytes())

    def upload(self, a, k: int = 0) -> None:
        """Copy a (float values) into entries k, k+1, ... along the first axis."""
        a = np.asarray(a)
        host = to_bf16_bits(a) if self.dtype == "bf16" else np.ascontiguousarray(a, dtype=np.float32)
        if host.nbytes + k * self.row_bytes() > self.nbytes:
            raise ValueError("upload out of range")
        L.check(L.load().shk_dev_copy(self.at(k), host.ctypes.data_as(C.c_void_p), host.nbytes, 0), None, "upload")

    def download(self) -> np.ndarray:
        host = np.empty(self.shape, dtype=np.uint16 if self.dtype == "bf16" else np.float32)
        L.check(L.load().shk_dev_copy(host.ctypes.data_as(C.c_void_p), self._h, self.nbytes, 1), None, "download")
        return from_bf16_bits(host) if self.dtype == "bf16" else host

    def copy_rows(self, dst_k: int, src: "DeviceArray", src_k: int, n: int = 1) -> None:
        if src.row_bytes() != self.row_bytes():
            raise ValueError("row sizes differ")
        L.check(L.load().shk_dev_copy(self.at(dst_k), src.at(src_k), n * self.row_bytes(), 2), None, "copy")

    @classmethod
    def from_numpy(cls, a, dtype: str = "f32") -> "DeviceArray":
        d = cls(np.shape(a), dtype)
        d.upload(a)
        return d


class UNet:
    """A network with buffers for up to ``max_batch`` images of
    ``max_h`` x ``max_w`` (multiples of 32 and 64)."""

    LOSSES = {"l2": 0, "energy": 1}

    def __init__(self, max_batch: int = 1, max_h: int = 32, max_w: int = 64, slope: float = 0.1,
                 seed: int | None = 0, params: np.ndarray | None = None, loss: str = "l2"):
        lib = L.load()
        self._h = C.c_void_p()
        cfg = L.ShkUnetCfg(int(max_batch), int(max_h), int(max_w), float(slope), self.LOSSES[loss])
        L.check_net(lib.shk_unet_create(C.byref(cfg), C.byref(self._h)), None, "shk_unet_create")
        weakref.finalize(self, _finalize_net, self._h)
        self.max_batch, self.max_h, self.max_w, self.slope = max_batch, max_h, max_w, slope
        self.n_params = int(lib.shk_unet_num_params())
        if params is not None:
            self.set_params(params)
        elif seed is not None:
            L.check_net(lib.shk_unet_init_params(self._h, C.c_uint64(seed)), self._h, "init")

    @property
    def handle(self):
        return self._h

    def _chk(self, rc, what):
        L.check_net(rc, self._h, what)

    def set_params(self, theta) -> None:
        a = np.ascontiguousarray(theta, dtype=np.float32)
        if a.size != self.n_params:
            raise L.ShaktiHIPError(f"expected {self.n_params} parameters, got {a.size}")
        self._chk(L.load().shk_unet_set_params(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "set")

    def get_params(self) -> np.ndarray:
        a = np.empty(self.n_params, dtype=np.float32)
        self._chk(L.load().shk_unet_get_params(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "get")
        return a

    def get_grads(self) -> np.ndarray:
        a = np.empty(self.n_params, dtype=np.float32)
        self._chk(L.load().shk_unet_get_grads(self._h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "grads")
        return a

    def forward(self, x: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
        """x: bf16 [B, H, W, 8] -> f32 [B, H, W] (a new buffer unless out is given)."""
        B, H, W, _ = x.shape
        if out is None:
            out = DeviceArray((B, H, W), "f32")
        self._chk(L.load().shk_unet_forward(self._h, x.ptr, B, H, W, out.ptr), "forward")
        return out

    def loss(self, t: DeviceArray, with_grad: bool = False) -> float:
        v = C.c_double()
        self._chk(L.load().shk_unet_loss(self._h, t.ptr, int(with_grad), C.byref(v)), "loss")
        return v.value

    def backward(self, dout: DeviceArray | None = None) -> None:
        self._chk(L.load().shk_unet_backward(self._h, dout.ptr if dout is not None else None), "backward")

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self._chk(L.load().shk_unet_adam(self._h, lr, beta1, beta2, eps), "adam")

    def train_step(self, x: DeviceArray, t: DeviceArray, lr: float) -> float:
        B, H, W, _ = x.shape
        v = C.c_double()
        self._chk(L.load().shk_unet_train_step(self._h, x.ptr, t.ptr, B, H, W, lr, C.byref(v)), "train_step")
        return v.value

    def tensor(self, name: str, B: int, H: int, W: int) -> np.ndarray:
        """An internal tensor of the last pass, as float32 [B, H>>l, W>>l, ch]
        (``out`` and ``dout`` without the channel axis)."""
        i = L.UNET_TENSOR_ID[name]
        _, ch, lvl = L.UNET_TENSORS[i]
        shape = (B, H >> lvl, W >> lvl, ch)
        a = np.empty(int(np.prod(shape)), dtype=np.float32)
        self._chk(L.load().shk_unet_download_tensor(self._h, i, a.ctypes.data_as(C.POINTER(C.c_float)), a.size),
                  f"tensor {name}")
        a = a.reshape(shape)
        return a[..., 0] if ch == 1 else a

    def save(self, path, **meta) -> None:
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.get_params(), slope=np.float32(self.slope),
                 meta=np.array(repr(meta)))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, max_batch: int = 1, max_h: int = 32, max_w: int = 64) -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            return cls(max_batch, max_h, max_w, slope=float(f["slope"]), seed=None, params=f["params"])


class PairRecorder:
    """Makes the State's elliptic solves fill device buffers with (network
    input, target) pairs: x bf16 [capacity, H8, W8, 8], t f32 [capacity, H8,
    W8], written from entry ``offset`` on, at most ``capacity`` pairs.

    kind "picard": one pair per Picard iteration (input: the residual at h;
    target: the scaled correction h_new - h).  kind "cg": one pair per
    learned CG iteration of a learned-preconditioner solve (input: that
    iteration's residual; target: its remaining error)."""

    KINDS = ("picard", "cg")

    def __init__(self, state, x: DeviceArray, t: DeviceArray, offset: int, capacity: int, kind: str = "picard"):
        if offset + capacity > x.shape[0] or offset + capacity > t.shape[0]:
            raise ValueError("recording past the end of the buffers")
        if kind not in self.KINDS:
            raise ValueError(f"unknown pair kind {kind!r}")
        self.state = state
        start = L.load().shk_record_pairs if kind == "picard" else L.load().shk_record_cg_pairs
        L.check(start(state.ctx, x.at(offset), t.at(offset), capacity), state.ctx, "record")

    @property
    def count(self) -> int:
        n = C.c_int32()
        L.check(L.load().shk_recorded_pairs(self.state.ctx, C.byref(n)), self.state.ctx, "recorded")
        return n.value

    def stop(self) -> int:
        n = self.count
        L.check(L.load().shk_record_pairs(self.state.ctx, None, None, 0), self.state.ctx, "record stop")
        return n


def glacier_variant(k: int, nx: int, ny: int, seed0: int = 1000):
    """The k-th training glacier: bed amplitude and correlation length vary
    with k so that bed features span about 10 to 80 cells at 512x128."""
    from . import synthetic as S
    amps = [20.0, 50.0, 100.0, 35.0]
    corr = [2000.0, 4000.0, 8000.0, 16000.0]
    g, raw = S.glacier(nx, ny, bed_amp=amps[k % 4], corr_len=corr[(k // 4) % 4], seed=seed0 + k)
    moulin = S.moulin_field(g, raw["mask"], n=20, seed=seed0 + k)
    return g, raw, moulin


def generate_pairs(n_glaciers: int = 64, nx: int = 512, ny: int = 128, per_glacier: int = 320,
                   max_steps: int = 200, seed0: int = 1000, rtol: float = 1e-10, log=print, kind: str = "picard",
                   precond: str = "multigrid", every: int = 16):
    """Run the plain solver (FP64 CG with the multigrid preconditioner, or
    ``precond="chebyshev"``: Chebyshev(4)) over seeded glaciers and record
    one pair per Picard iteration (kind "picard"), or one pair per
    ``every``-th CG iteration (kind "cg": the learned-preconditioner solve
    with a network whose output is zero, which takes plain PCG's steps).
    Returns the device buffers (x, t) and the number of pairs of each
    glacier (glacier k holds entries k*per_glacier ...)."""
    from . import api as H
    from . import synthetic as S
    H8, W8 = canvas_shape(nx, ny)
    cap = n_glaciers * per_glacier
    x = DeviceArray((cap, H8, W8, 8), "bf16")
    t = DeviceArray((cap, H8, W8), "f32")
    counts = []
    kw = dict(multigrid=True) if precond == "multigrid" else dict(chebyshev_degree=4)
    if kind == "cg":
        zero = UNet(1, H8, W8, seed=None)
        zero.set_params(np.zeros(zero.n_params, dtype=np.float32))
        kw.update(form="learned", network=zero, learned_precond_every=every, stall_ratio=1e3)
    t0 = time.perf_counter()
    for k in range(n_glaciers):
        g, raw, moulin = glacier_variant(k, nx, ny, seed0)
        p = H.ModelParameters()
        sl = H.RegularizedCoulombSlidingLaw(0.5)
        s = H.State(g, p)
        H.set_initial_conditions(s, g, p, sl, **raw)
        ls = H.CGIterativeSolver(g, H.HIPMatrixFreeLinearSystem, atol=0.0, rtol=rtol, check_every=16, **kw)
        ps = H.PicardSolver(500, 1e-6, ls, g)
        mi = H.SeasonalMoulinMeltInput(moulin, **S.SEASONAL_C2)
        sim = H.Simulation(g, s, max_steps, 3600.0, p, "implicit", [], mi, sl, ps=ps, k_face_choice="harmonic")
        rec = PairRecorder(s, x, t, k * per_glacier, per_glacier, kind)
        steps = 0
        while rec.count < per_glacier and steps < max_steps:
            H.step(sim)
            sim.total_time += sim.dt
            steps += 1
        counts.append(rec.stop())
        if log and (k % 8 == 7 or k == n_glaciers - 1):
            log(f"[pairs] glacier {k + 1}/{n_glaciers}: {counts[-1]} pairs in {steps} steps "
                f"({time.perf_counter() - t0:.1f} s)")
    return x, t, counts


def _fill_batch(xb: DeviceArray, tb: DeviceArray, x: DeviceArray, t: DeviceArray, sel) -> None:
    for i, k in enumerate(sel):
        xb.copy_rows(i, x, int(k))
        tb.copy_rows(i, t, int(k))


def train(net: UNet, x: DeviceArray, t: DeviceArray, idx_train, iters: int = 2000, batch: int = 16,
          lr: float = 1e-3, lr_min: float = 1e-5, seed: int = 0, idx_val=None, log=print, log_every: int = 200):
    """Adam with a cosine learning-rate decay on random batches of the
    recorded pairs; returns the training-loss history and the validation
    losses (every log_every iterations)."""
    rng = np.random.default_rng(seed)
    idx_train = np.asarray(idx_train)
    _, H8, W8, _ = x.shape
    xb = DeviceArray((batch, H8, W8, 8), "bf16")
    tb = DeviceArray((batch, H8, W8), "f32")
    hist, val = [], []
    t0 = time.perf_counter()
    for it in range(iters):
        _fill_batch(xb, tb, x, t, idx_train[rng.integers(0, len(idx_train), batch)])
        lr_it = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / max(1, iters - 1)))
        hist.append(net.train_step(xb, tb, lr_it))
        if (it + 1) % log_every == 0 or it == iters - 1:
            v = evaluate(net, x, t, idx_val, batch) if idx_val is not None else float("nan")
            val.append((it + 1, v))
            if log:
                log(f"[train] iter {it + 1}/{iters}: loss {np.mean(hist[-log_every:]):.4f}, val {v:.4f}, "
                    f"lr {lr_it:.2e} ({time.perf_counter() - t0:.1f} s)")
    return hist, val


def evaluate(net: UNet, x: DeviceArray, t: DeviceArray, idx, batch: int = 16) -> float:
    """Mean scale-free loss over the pairs idx."""
    idx = np.asarray(idx)
    _, H8, W8, _ = x.shape
    tot, n = 0.0, 0
    for i in range(0, len(idx), batch):
        sel = idx[i:i + batch]
        xb = DeviceArray((len(sel), H8, W8, 8), "bf16")
        tb = DeviceArray((len(sel), H8, W8), "f32")
        _fill_batch(xb, tb, x, t, sel)
        net.forward(xb)
        tot += net.loss(tb) * len(sel)
        n += len(sel)
    return tot / max(1, n)
