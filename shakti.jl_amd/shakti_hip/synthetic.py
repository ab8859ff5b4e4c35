"""Deterministic synthetic glaciers for the benchmark and test workloads.

Every choice is fixed here so that the GPU run and the CPU checker receive
bit-identical inputs from the same seed:

* grid: ``Grid(nx, ny, 100e3, 20e3)`` (a 100 km x 20 km domain);
* bed ``zb``: a zero-mean Gaussian random field -- white noise from numpy's
  PCG64 generator seeded with ``seed``, smoothed by a separable Gaussian kernel
  of standard deviation ``corr_len/sqrt(2)`` (so the field's covariance is
  ``exp(-r^2 / (2 corr_len^2))``), periodic boundaries, then shifted to zero
  mean and scaled to standard deviation ``bed_amp``;
  optionally a trough: a Gaussian channel of full width ``trough_width`` and
  depth ``trough_depth`` along x at a seeded y position;
* surface ``zs(x) = 6 (sqrt(x + 5000) - sqrt(5000)) + 1`` m;
* mask: all GROUNDED except the ``ix = 0`` column, which is OCEAN;
* ``b0 = 0.005`` m, ``A_visc = 2.5e-25``, ``ub_x = 1e-6`` m/s, ``ub_y = 0``,
  ``G = 0.05`` W/m^2, ``ieb = 1e-8`` m/s, ``taub = 0``;
* moulins: ``n`` distinct GROUNDED interior cells chosen with the seed, each
  with ``Q ~ U[qmin, qmax]`` m^3/s, giving ``ieb += Q / (dx dy)``.

``generate_on_device`` builds a glacier of the same description on the GPU
(shk_generate_glacier) from a counter-based stream (SplitMix64) instead of
numpy's PCG64 and a truncated direct convolution instead of the FFT, so its
fields differ from ``glacier``'s for the same seed; a CPU restatement
reproduces them bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .api import GROUNDED, OCEAN, Grid


def _smooth_periodic(w: np.ndarray, sx: float, sy: float) -> np.ndarray:
    """Separable Gaussian smoothing with periodic wrap, via FFT."""
    nx, ny = w.shape
    kx = np.fft.fftfreq(nx)
    ky = np.fft.fftfreq(ny)
    gx = np.exp(-2.0 * (math.pi * kx * sx) ** 2)
    gy = np.exp(-2.0 * (math.pi * ky * sy) ** 2)
    W = np.fft.fft2(w)
    return np.real(np.fft.ifft2(W * gx[:, None] * gy[None, :]))


def bed_field(g: Grid, bed_amp: float, corr_len: float = 2000.0, seed: int = 0) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((g.nx, g.ny))
    sk = corr_len / math.sqrt(2.0)
    f = _smooth_periodic(w, sk / g.dx, sk / g.dy)
    f -= f.mean()
    sd = f.std()
    if sd > 0:
        f *= bed_amp / sd
    return np.asfortranarray(f)


def surface(g: Grid) -> np.ndarray:
    zs = 6.0 * (np.sqrt(g.x + 5000.0) - math.sqrt(5000.0)) + 1.0
    return np.asfortranarray(np.repeat(zs.reshape(g.nx, 1), g.ny, axis=1))


def moulin_field(g: Grid, mask: np.ndarray, n: int = 20, qmin: float = 1.0, qmax: float = 10.0,
                 seed: int = 0) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    cand = [(ix, iy) for iy in range(1, g.ny - 1) for ix in range(1, g.nx - 1) if mask[ix, iy] == GROUNDED]
    pick = rng.choice(len(cand), size=min(n, len(cand)), replace=False)
    q = rng.uniform(qmin, qmax, size=len(pick))
    m = np.zeros((g.nx, g.ny), order="F")
    for k, j in enumerate(pick):
        ix, iy = cand[int(j)]
        m[ix, iy] += q[k] / (g.dx * g.dy)
    return m


def glacier(nx: int, ny: int, bed_amp: float = 10.0, corr_len: float = 2000.0, seed: int = 0,
            lx: float = 100e3, ly: float = 20e3, trough_width: float = 0.0, trough_depth: float = 0.0,
            b0: float = 0.005, A: float = 2.5e-25, ub: float = 1e-6, G: float = 0.05, ieb: float = 1e-8):
    """Returns (grid, raw inputs dict for set_initial_conditions)."""
    g = Grid(nx, ny, lx, ly)
    zb = bed_field(g, bed_amp, corr_len, seed)
    if trough_width > 0 and trough_depth != 0:
        rng = np.random.Generator(np.random.PCG64(seed + 104729))
        y0 = rng.uniform(0.25, 0.75) * ly
        sig = trough_width / 2.0
        zb = zb - trough_depth * np.exp(-0.5 * ((g.y[None, :] - y0) / sig) ** 2)
        zb = np.asfortranarray(zb)
    mask = np.full((nx, ny), GROUNDED, order="F")
    mask[0, :] = OCEAN
    raw = dict(
        mask=mask,
        A_visc=np.full((nx, ny), A, order="F"),
        zb=zb,
        zs=surface(g),
        b=np.full((nx, ny), b0, order="F"),
        G=np.full((nx, ny), G, order="F"),
        ub_x=np.full((nx + 1, ny), ub, order="F"),
        ub_y=np.zeros((nx, ny + 1), order="F"),
        ieb=np.full((nx, ny), ieb, order="F"),
        taub_x=np.zeros((nx + 1, ny), order="F"),
        taub_y=np.zeros((nx, ny + 1), order="F"),
    )
    return g, raw


def generate_on_device(s, seed: int = 0, bed_amp: float = 10.0, corr_len: float = 2000.0,
                       trough_width: float = 0.0, trough_depth: float = 0.0, b0: float = 0.005, A: float = 2.5e-25,
                       ub: float = 1e-6, G: float = 0.05, ieb: float = 1e-8, n_moulins: int = 0, qmin: float = 1.0,
                       qmax: float = 10.0) -> None:
    """Writes the raw inputs of set_initial_conditions (and, with n_moulins,
    the moulin field) into the device State ``s`` (shk_generate_glacier);
    follow with api.set_initial_conditions_on_device."""
    import ctypes as C
    from . import _lib as L
    spec = L.ShkGlacierSpec(int(seed), bed_amp, corr_len, trough_width, trough_depth, b0, A, ub, G, ieb,
                            int(n_moulins), 0, qmin, qmax)
    L.check(L.load().shk_generate_glacier(s.ctx, C.byref(spec)), s.ctx, "shk_generate_glacier")


def seasonal_c2(t: float) -> float:
    """The uniform forcing of the 1024x256 workload: 1e-8 m/s baseline plus a
    1e-7 m/s one-day sinusoid, i.e. SeasonalMeltInput(t_start=0, t_end=1,
    offset=6e-8*86400, amplitude=5e-8*86400, seconds_per_year=86400)."""
    spy = 86400.0
    yf = math.fmod(t / spy, 1.0)
    return (6e-8 * spy - 5e-8 * spy * math.cos(2 * math.pi * yf)) / spy


SEASONAL_C2 = dict(t_start=0.0, t_end=1.0, offset=6e-8 * 86400, amplitude=5e-8 * 86400, i_min=0.0,
                   seconds_per_year=86400.0)
