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
  finest level takes the two-column legs; coarsening uses the FP32 values;
* bil_from: the transfers from level bil_from down are cell-centred linear
  ones instead: P takes a fine cell 3/4 of its own coarse cell's value plus
  1/4 of the neighbour on its side, in x and in y (9/16, 3/16, 3/16, 1/16),
  a neighbour beyond the grid folding onto the cell's own (weight 1);
  restriction is P^T and the coarse-correction scale is scale_bil; the
  coarser level's operator is the Galerkin product P^T A P lumped back to the
  5-point flux form: c = P^T c, and each fine face of coupling a between
  cells i, j contributes a * (P_i - P_j) (P_i - P_j)^T, whose links are laid
  onto the lattice paths between their cells (half on each of the two paths
  of a diagonal link).  For an x-face that leaves the couplings
    ec(I, J) = sum_y wy(y, J) [a(2I+1, y) / 4 + (a(2I, y) + a(2I+2, y)) / 16]
  (faces beyond the grid taken as zero; wy the restriction weights in y) and
  nothing on the y-faces; y-faces likewise.

* nu2_fine: the finest level takes two sweeps each way too (the library's
  choice where the finest level takes the BF16 two-column legs);
* nu2: levels (below the finest, reached by the linear transfers) that take
  two sweeps each way: down-leg red, black, red, black (the first red cells
  from zero, every later half-sweep x += omega (b - A x) / D on its colour),
  up-leg black, red, black, red.  `library_nu2` is the library's rule for
  which levels do: those that take the two-row legs (even width, more than
  4096 cells) and keep two rows per wave with the wider halo, 116 owned
  columns per wave: ny * ceil(nx / 116) // 4096 <= 2.

Arrays are indexed [ix, iy] with shape (nx, ny); everything is float32 (the
linear transfers and the coarsening in float64, rounded once).
"""
import numpy as np

F32 = np.float32


def lin_weights(f):
    """The cell-centred linear transfer on a 1D level of f cells, per fine index
    x: (I0, w0, I1, w1), the two coarse cells and their weights; a neighbour
    outside [0, (f + 1) // 2) folds onto I0 (w0 = 1, w1 = 0)."""
    c = (f + 1) // 2
    x = np.arange(f)
    I0 = x // 2
    I1 = np.where(x % 2 == 0, I0 - 1, I0 + 1)
    out = (I1 < 0) | (I1 > c - 1)
    return I0, np.where(out, 1.0, 0.75), np.clip(I1, 0, c - 1), np.where(out, 0.0, 0.25)


def _restrict1(a, axis):
    f = a.shape[axis]
    I0, w0, I1, w1 = lin_weights(f)
    a = np.moveaxis(np.asarray(a, np.float64), axis, 0)
    out = np.zeros(((f + 1) // 2,) + a.shape[1:])
    np.add.at(out, I0, w0[:, None] * a)
    np.add.at(out, I1, w1[:, None] * a)
    return np.moveaxis(out, 0, axis)


def _prolong1(a, f, axis):
    I0, w0, I1, w1 = lin_weights(f)
    a = np.moveaxis(np.asarray(a, np.float64), axis, 0)
    return np.moveaxis(w0[:, None] * a[I0] + w1[:, None] * a[I1], 0, axis)


def restrict_lin(r):
    """P^T r (float64)."""
    return _restrict1(_restrict1(r, 0), 1)


def prolong_lin(xc, fx, fy):
    """P xc onto an fx x fy level (float64)."""
    return _prolong1(_prolong1(xc, fx, 0), fy, 1)


def coarsen_lin(c, e, n):
    """The lumped Galerkin operator of the linear transfers (float64 in, FP32 out)."""
    fx, fy = c.shape
    cx, cy = (fx + 1) // 2, (fy + 1) // 2
    ep = np.zeros((2 * cx + 2, fy))
    ep[:fx - 1] = np.asarray(e, np.float64)[:fx - 1]
    hx = 0.25 * ep[1:2 * cx:2] + 0.0625 * (ep[0:2 * cx:2] + ep[2:2 * cx + 2:2])
    hx[-1] = 0.0
    npd = np.zeros((fx, 2 * cy + 2))
    npd[:, :fy - 1] = np.asarray(n, np.float64)[:, :fy - 1]
    hy = 0.25 * npd[:, 1:2 * cy:2] + 0.0625 * (npd[:, 0:2 * cy:2] + npd[:, 2:2 * cy + 2:2])
    hy[:, -1] = 0.0
    return restrict_lin(c).astype(F32), _restrict1(hx, 1).astype(F32), _restrict1(hy, 0).astype(F32)


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


def library_nu2(nx, ny):
    """Whether the library's level of nx x ny cells (with the linear transfers)
    smooths twice each way."""
    if nx * ny <= 4096 or nx % 2:
        return False
    strips = -(-nx // 116)
    return (ny * strips) // 4096 <= 2


class VCycle:
    def __init__(self, aP, aE, aN, omega=1.2, scale=1.9, bf16_level0=False, bil_from=None, scale_bil=0.8,
                 nu2=None, nu2_fine=False):
        self.omega, self.scale, self.scale_bil = F32(omega), F32(scale), F32(scale_bil)
        c, e, n, D = fine_level(aP, aE, aN)
        self.levels = [(c, e, n, D)]
        self.lin = []  # per level: the transfer to the next level is linear
        while True:
            c, e, n, _ = self.levels[-1]
            if c.size <= 64 and len(self.levels) > 1:
                break
            lin = bil_from is not None and len(self.levels) - 1 >= bil_from
            cc, ec, nc = (coarsen_lin if lin else coarsen)(c, e, n)
            self.lin.append(lin)
            self.levels.append((cc, ec, nc, diag(cc, ec, nc)))
        if bf16_level0:
            self.levels[0] = tuple(bf16(a) for a in self.levels[0])
        c, e, n, _ = self.levels[-1]
        self.inv = dense_inverse(c, e, n)
        self.red = [((np.arange(L[0].shape[0])[:, None] + np.arange(L[0].shape[1])[None, :]) % 2) == 0
                    for L in self.levels]
        # sweeps each way per level: two where nu2 (a rule on the level's
        # shape) says so on a level below the finest with the linear transfers
        self.nu = [2 if (nu2 is not None and 0 < l < len(self.lin) and self.lin[l] and nu2(*L[0].shape)) else 1
                   for l, L in enumerate(self.levels)]
        if nu2_fine and len(self.levels) > 1:
            self.nu[0] = 2

    def _cycle(self, b, l):
        c, e, n, D = self.levels[l]
        if l == len(self.levels) - 1:
            nx, ny = b.shape
            flat = b.reshape(-1, order="F")
            return (self.inv @ flat).reshape((nx, ny), order="F").astype(F32)
        om, red = self.omega, self.red[l]
        blk = ~red

        def relax(u, mask):
            t = u + om * (b - flux_apply(c, e, n, u)) / D
            u[mask] = t[mask]

        x = np.zeros_like(b)
        x[red] = om * b[red] / D[red]
        s = b + neighbour_sum(e, n, x)
        x[blk] = om * s[blk] / D[blk]
        for _ in range(self.nu[l] - 1):
            relax(x, red)
            relax(x, blk)
        r = b - flux_apply(c, e, n, x)
        fx, fy = b.shape
        cx, cy = self.levels[l + 1][0].shape
        if self.lin[l]:
            xc = self._cycle(restrict_lin(r).astype(F32), l + 1)
            u = x + (self.scale_bil * prolong_lin(xc, fx, fy)).astype(F32)
        else:
            rp = np.zeros((2 * cx, 2 * cy), F32)
            rp[:fx, :fy] = r
            rc = (rp[0::2, 0::2] + rp[1::2, 0::2]) + (rp[0::2, 1::2] + rp[1::2, 1::2])
            xc = self._cycle(rc.astype(F32), l + 1)
            u = x + self.scale * np.repeat(np.repeat(xc, 2, 0), 2, 1)[:fx, :fy]
        for _ in range(self.nu[l]):
            relax(u, blk)
            relax(u, red)
        return u

    def __call__(self, r):
        """z = M r for an FP64 (nx, ny) array; FP32 inside, FP64 out."""
        return self._cycle(np.asarray(r, dtype=F32), 0).astype(np.float64)
