"""NumPy restatement of the library's multigrid preconditioner (test helper).

The V-cycle is this build's own preconditioner (the reference has Jacobi,
Chebyshev and CPU-side AMG), so it is restated here from its definition, not
from a reference file:

* hierarchy: the FP32 flux form of the FP64 system, c = diagonal minus the kept
  couplings, e / n the east / north couplings, D the FP32 diagonal; each coarser
  level sums 2 x 2 blocks (c over the block, e and n over the fine faces on the
  block's east and north sides) until at most 64 cells remain (at least once);
  its D is c + e + n + e_west + n_south; the coarsest level is solved with its
  dense inverse (formed in FP64, kept in FP32);
* smoothing: symmetric red-black Gauss-Seidel with over-relaxation omega (red
  cells ix + iy even): down-leg from x = 0, red then black; up-leg after the
  coarse correction u = x + scale P xc, black then red;
* bf16_level0: the finest level's smoothing and residual use its c, e, n, D
  rounded to BF16 (round to nearest even), as the library does where the
  finest level takes the two-column legs; coarsening uses the FP32 values.

Arrays are indexed [ix, iy] with shape (nx, ny); everything is float32.
"""
import numpy as np

F32 = np.float32


def flux_apply(c, e, n, u):
    """A u in the flux form: c u + sum over faces a_f (u - u_f), terms E, W, N, S."""
    y = c * u
    y[:-1] += e[:-1] * (u[:-1] - u[1:])
    y[1:] += e[:-1] * (u[1:] - u[:-1])
    y[:, :-1] += n[:, :-1] * (u[:, :-1] - u[:, 1:])
    y[:, 1:] += n[:, :-1] * (u[:, 1:] - u[:, :-1])
    return y


def neighbour_sum(e, n, x):
    """sum over faces a_f x_f (W, E, N, S), absent faces contributing nothing."""
    s = np.zeros_like(x)
    s[1:] += e[:-1] * x[:-1]
    s[:-1] += e[:-1] * x[1:]
    s[:, :-1] += n[:, :-1] * x[:, 1:]
    s[:, 1:] += n[:, :-1] * x[:, :-1]
    return s


def bf16(a):
    """FP32 -> BF16 -> FP32, round to nearest even."""
    u = np.asarray(a, dtype=F32).view(np.uint32)
    u = ((u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)).astype(np.uint32)
    return u.view(F32)


def fine_level(aP, aE, aN):
    """FP32 flux form of the FP64 system (symmetric face layout, [ix, iy])."""
    aP, aE, aN = (np.asarray(a, dtype=np.float64) for a in (aP, aE, aN))
    E = aE.copy()
    E[-1, :] = 0.0
    N = aN.copy()
    N[:, -1] = 0.0
    W = np.zeros_like(E)
    W[1:] = E[:-1]
    S = np.zeros_like(N)
    S[:, 1:] = N[:, :-1]
    c = aP - (E + W + N + S)
    return c.astype(F32), E.astype(F32), N.astype(F32), aP.astype(F32)


def coarsen(c, e, n):
    """2 x 2 aggregation; a block's cells summed as ((c00 + c10) + c01) + c11,
    cells beyond an odd edge counting as zero."""
    fx, fy = c.shape
    cx, cy = (fx + 1) // 2, (fy + 1) // 2

    def pad(a):
        p = np.zeros((2 * cx, 2 * cy), F32)
        p[:fx, :fy] = a
        return p

    cp, ep, np_ = pad(c), pad(e), pad(n)
    C = ((cp[0::2, 0::2] + cp[1::2, 0::2]) + cp[0::2, 1::2]) + cp[1::2, 1::2]
    Ec = ep[1::2, 0::2] + ep[1::2, 1::2]  # faces on the block's east side
    Ec[-1, :] = 0.0
    Nc = np_[0::2, 1::2] + np_[1::2, 1::2]  # faces on its north side
    Nc[:, -1] = 0.0
    return C.astype(F32), Ec.astype(F32), Nc.astype(F32)


def diag(c, e, n):
    d = c + e + n
    d[1:] += e[:-1]
    d[:, 1:] += n[:, :-1]
    return d


def dense_inverse(c, e, n):
    nx, ny = c.shape
    m = nx * ny
    A = np.zeros((m, m))
    idx = lambda ix, iy: ix + iy * nx  # the device's cell order
    for iy in range(ny):
        for ix in range(nx):
            k = idx(ix, iy)
            A[k, k] += float(c[ix, iy])
            if ix < nx - 1:
                a = float(e[ix, iy])
                A[k, k] += a
                A[idx(ix + 1, iy), idx(ix + 1, iy)] += a
                A[k, idx(ix + 1, iy)] -= a
                A[idx(ix + 1, iy), k] -= a
            if iy < ny - 1:
                a = float(n[ix, iy])
                A[k, k] += a
                A[idx(ix, iy + 1), idx(ix, iy + 1)] += a
                A[k, idx(ix, iy + 1)] -= a
                A[idx(ix, iy + 1), k] -= a
    return np.linalg.inv(A).astype(F32)


class VCycle:
    def __init__(self, aP, aE, aN, omega=1.2, scale=1.8, bf16_level0=False):
        self.omega, self.scale = F32(omega), F32(scale)
        c, e, n, D = fine_level(aP, aE, aN)
        self.levels = [(c, e, n, D)]
        while True:
            c, e, n, _ = self.levels[-1]
            if c.size <= 64 and len(self.levels) > 1:
                break
            cc, ec, nc = coarsen(c, e, n)
            self.levels.append((cc, ec, nc, diag(cc, ec, nc)))
        if bf16_level0:
            self.levels[0] = tuple(bf16(a) for a in self.levels[0])
        c, e, n, _ = self.levels[-1]
        self.inv = dense_inverse(c, e, n)
        self.red = [((np.arange(L[0].shape[0])[:, None] + np.arange(L[0].shape[1])[None, :]) % 2) == 0
                    for L in self.levels]

    def _cycle(self, b, l):
        c, e, n, D = self.levels[l]
        if l == len(self.levels) - 1:
            nx, ny = b.shape
            flat = b.reshape(-1, order="F")
            return (self.inv @ flat).reshape((nx, ny), order="F").astype(F32)
        om, red = self.omega, self.red[l]
        blk = ~red
        x = np.zeros_like(b)
        x[red] = om * b[red] / D[red]
        s = b + neighbour_sum(e, n, x)
        x[blk] = om * s[blk] / D[blk]
        r = b - flux_apply(c, e, n, x)
        fx, fy = b.shape
        cx, cy = self.levels[l + 1][0].shape
        rp = np.zeros((2 * cx, 2 * cy), F32)
        rp[:fx, :fy] = r
        rc = (rp[0::2, 0::2] + rp[1::2, 0::2]) + (rp[0::2, 1::2] + rp[1::2, 1::2])
        xc = self._cycle(rc.astype(F32), l + 1)
        u = x + self.scale * np.repeat(np.repeat(xc, 2, 0), 2, 1)[:fx, :fy]
        for mask in (blk, red):
            t = u + om * (b - flux_apply(c, e, n, u)) / D
            u[mask] = t[mask]
        return u

    def __call__(self, r):
        """z = M r for an FP64 (nx, ny) array; FP32 inside, FP64 out."""
        return self._cycle(np.asarray(r, dtype=F32), 0).astype(np.float64)
