"""ORACLE -- test infrastructure only.

ctypes front-end of ``libshakti_oracle.so``, the plain-C FP64 restatement of
Shakti.jl's hydraulic-head path (see ``shakti_oracle.c`` for the per-function
citations).  Only the tests, the smoke check and the bench's CPU-baseline leg
import this module; the product library never does.

Arrays follow the reference layout (src/state.jl:13-57): Fortran-ordered numpy
arrays indexed ``a[ix, iy]`` with shapes (nx, ny) for centre fields,
(nx+1, ny) for x-face fields and (nx, ny+1) for y-face fields.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
from dataclasses import dataclass, fields

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libshakti_oracle.so"

CENTER_FIELDS = ["h", "pw", "po", "b", "beta", "abs_ub", "mdot", "shear", "potential", "sensible", "Re", "K", "G",
                 "zb", "zs", "H", "ieb", "lambda", "A_visc", "N", "mask"]
XFACE_FIELDS = ["dhdx", "q_x", "Re_x", "b_x", "ub_x", "taub_x", "dpwdx", "valid_x"]
YFACE_FIELDS = ["dhdy", "q_y", "Re_y", "b_y", "ub_y", "taub_y", "dpwdy", "valid_y"]
ALL_FIELDS = CENTER_FIELDS + XFACE_FIELDS + YFACE_FIELDS

GROUNDED, OCEAN, LAND, OTHER_BASIN = 0.0, 1.0, 2.0, 3.0


def build(force: bool = False) -> pathlib.Path:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    import subprocess
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-B" if force else "-s", "-C", str(_HERE)], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("lx", C.c_double), ("ly", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("dx2", C.c_double), ("dy2", C.c_double)]


class _Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ["rho_w", "rho_sw", "rho_i", "g", "nu", "n", "omega", "L", "br", "lr", "ct",
                                          "cw", "p_atm", "b_min", "b_max", "e_v"]]


class _Model(C.Structure):
    _fields_ = [("kfs", C.c_int32), ("gap_scheme", C.c_int32), ("sliding_law", C.c_int32),
                ("sliding_C", C.c_double), ("v0", C.c_double), ("Cx", C.POINTER(C.c_double)),
                ("Cy", C.POINTER(C.c_double))]


class _GlacierSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64)] + [(n, C.c_double) for n in ["bed_amp", "corr_len", "trough_width",
                                                                    "trough_depth", "b0", "A_visc", "ub", "G",
                                                                    "ieb"]] + \
        [("n_moulins", C.c_int32), ("reserved", C.c_int32), ("q_min", C.c_double), ("q_max", C.c_double)]


class _State(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ALL_FIELDS]


class _CgCfg(C.Structure):
    _fields_ = [("precond", C.c_int32), ("cheb_degree", C.c_int32), ("cheb_nsteps", C.c_int32),
                ("itmax", C.c_int32), ("atol", C.c_double), ("rtol", C.c_double), ("lambda_min", C.c_double),
                ("lambda_max", C.c_double), ("gershgorin", C.c_int32), ("pad", C.c_int32)]


class CgStats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("solved", C.c_int32), ("zero_curvature", C.c_int32),
                ("bounds_warning", C.c_int32), ("rnorm0", C.c_double), ("rnorm", C.c_double),
                ("lambda_min", C.c_double), ("lambda_max", C.c_double), ("breakdown", C.c_int32),
                ("pad", C.c_int32)]


class _Op(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32)] + [(n, C.POINTER(C.c_double)) for n in
                                                          ["aP", "aE", "aW", "aN", "aS", "d"]]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = C.CDLL(str(_LIB_PATH))
        dp = C.POINTER(C.c_double)
        L.orc_make_grid.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.POINTER(_Grid)]
        L.orc_default_params.argtypes = [C.POINTER(_Params)]
        for nm in ["orc_face_masks", "orc_apply_mask_to_sliding"]:
            getattr(L, nm).argtypes = [C.POINTER(_Grid), C.POINTER(_State)]
        for nm in ["orc_set_initial_conditions", "orc_update_derived", "orc_test_refresh_lagged"]:
            getattr(L, nm).argtypes = [C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_Model), C.POINTER(_State)]
        L.orc_lagged_q_then_Re.argtypes = [C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_State)]
        L.orc_step_b.argtypes = [C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_Model), C.POINTER(_State),
                                 C.c_double]
        L.orc_assemble_mfls.argtypes = [C.POINTER(_Grid), C.POINTER(_Params), C.c_int32, C.POINTER(_State),
                                        C.c_int32, C.c_double] + [dp] * 6
        L.orc_assemble_sals.argtypes = [C.POINTER(_Grid), C.POINTER(_Params), C.c_int32, C.POINTER(_State),
                                        C.c_int32, C.c_double, C.POINTER(C.c_int64), dp, dp]
        L.orc_matvec.argtypes = [C.c_int32, C.c_int32] + [dp] * 7
        L.orc_dot.argtypes = [C.c_size_t, dp, dp]
        L.orc_dot.restype = C.c_double
        L.orc_tridiag_extreme_eigs.argtypes = [C.c_int32, dp, dp, dp, dp]
        L.orc_estimate_eigenvalue_bounds.argtypes = [C.POINTER(_Op), dp, C.c_int32, dp, dp]
        L.orc_update_chebyshev_bounds.argtypes = [C.POINTER(_Op), dp, C.c_int32, dp, dp]
        L.orc_cheb_apply.argtypes = [C.POINTER(_Op), C.c_int32, C.c_double, C.c_double, dp, dp, dp, dp, dp]
        L.orc_cg_solve.argtypes = [C.POINTER(_Op), C.POINTER(_CgCfg), dp, dp, C.POINTER(CgStats)]
        L.orc_gershgorin.argtypes = [C.POINTER(_Op)]
        L.orc_estimate_eigenvalue_bounds_d.argtypes = [C.POINTER(_Op), dp, C.c_int32, dp, dp]
        L.orc_gershgorin.restype = C.c_double
        L.orc_cholesky_solve.argtypes = [C.c_int32, C.c_int32] + [dp] * 7
        L.orc_solver_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(_CgCfg)]
        L.orc_solver_create.restype = C.c_void_p
        L.orc_solver_destroy.argtypes = [C.c_void_p]
        L.orc_solver_bounds.argtypes = [C.c_void_p, dp, dp]
        L.orc_solve_linear.argtypes = [C.c_void_p, C.POINTER(_Grid), C.POINTER(_Params), C.c_int32,
                                       C.POINTER(_State), C.c_int32, C.c_double, C.POINTER(CgStats)]
        L.orc_picard_loop.argtypes = [C.c_void_p, C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_Model),
                                      C.POINTER(_State), C.c_int32, C.c_double, C.c_int32, C.c_double,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.orc_picard_loop.restype = C.c_int32
        L.orc_parabolic_step.argtypes = [C.c_void_p, C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_Model),
                                         C.POINTER(_State), C.c_double, C.POINTER(C.c_int64)]
        L.orc_seasonal_ieb.argtypes = [C.c_double] * 7
        L.orc_seasonal_ieb.restype = C.c_double
        L.orc_step.argtypes = [C.c_void_p, C.POINTER(_Grid), C.POINTER(_Params), C.POINTER(_Model),
                               C.POINTER(_State), C.c_double, C.c_int32, C.c_double, dp, C.c_int32, C.c_double,
                               C.c_int32, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.orc_step.restype = C.c_int32
        L.orc_num_threads.restype = C.c_int
        L.orc_stagger_c.argtypes = [C.POINTER(_Grid), dp, dp, dp]
        L.orc_generate_glacier.argtypes = [C.POINTER(_Grid), C.POINTER(_GlacierSpec), C.POINTER(_State), dp]
        L.orc_generate_glacier.restype = C.c_int
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous, "oracle arrays must be Fortran-ordered float64"
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- data model

@dataclass
class Grid:
    """src/grid.jl:26-35 -- cell centres span [0, lx] x [0, ly]."""
    nx: int
    ny: int
    lx: float
    ly: float

    def __post_init__(self):
        self.dx = self.lx / (self.nx - 1)
        self.dy = self.ly / (self.ny - 1)
        self.dx2 = self.dx ** 2
        self.dy2 = self.dy ** 2
        self.x = np.arange(self.nx) * self.dx
        self.y = np.arange(self.ny) * self.dy

    def c(self) -> _Grid:
        g = _Grid()
        lib().orc_make_grid(self.nx, self.ny, self.lx, self.ly, C.byref(g))
        return g


@dataclass
class ModelParameters:
    """src/model_parameters.jl:55-84 defaults."""
    rho_w: float = 1000.0
    rho_sw: float = 1027.0
    rho_i: float = 910.0
    g: float = 9.81
    nu: float = 1.787e-6
    n: float = 3.0
    omega: float = 1e-4
    L: float = 334e3
    br: float = 0.05
    lr: float = 2.0
    ct: float = 7.5e-8
    cw: float = 4.22e3
    p_atm: float = 0.0
    b_min: float = 0.0
    b_max: float = float("inf")
    e_v: float = 0.0

    def c(self) -> _Params:
        return _Params(*[getattr(self, f.name) for f in fields(self)])


SLIDING_LAWS = {"regularized_coulomb": 0, "prescribed": 1, "linear": 2, "regularized_coulomb_field": 3,
                "regularized_coulomb_v0": 4}


def stagger_c(g: Grid, Cc) -> tuple[np.ndarray, np.ndarray]:
    """C at the cell centres (an (nx, ny) array or a scalar) staggered onto the x- and y-faces as
    LinearSlidingLaw(g, C), RegularizedCoulombFieldSlidingLaw(g, C) and
    RegularizedCoulombV0SlidingLaw(g, C, v0) stagger it (src/sliding_law.jl:73-90, 284-301, 436-452);
    the linear law squares the face values where it uses them."""
    cc = F(np.broadcast_to(np.asarray(Cc, dtype=np.float64), (g.nx, g.ny)))
    Cx = np.zeros((g.nx + 1, g.ny), order="F")
    Cy = np.zeros((g.nx, g.ny + 1), order="F")
    lib().orc_stagger_c(C.byref(g.c()), _p(cc), _p(Cx), _p(Cy))
    return Cx, Cy


@dataclass
class Model:
    """The scheme choices of a Simulation (src/simulation.jl:105-146).  ``sliding_law`` is a code of
    SLIDING_LAWS; the field laws (linear, regularized_coulomb_field, regularized_coulomb_v0) take their
    face coefficients Cx, Cy from ``stagger_c`` (``Model.field_law`` builds one)."""
    kfs: int = 0          # 0 Arithmetic, 1 Harmonic
    gap_scheme: int = 0   # 0 implicit, 1 explicit
    sliding_law: int = 0  # SLIDING_LAWS
    sliding_C: float = 0.5
    v0: float = 0.0
    Cx: np.ndarray | None = None
    Cy: np.ndarray | None = None

    @classmethod
    def field_law(cls, g: Grid, law: str, Cc, v0: float = 0.0, **kw) -> "Model":
        Cx, Cy = stagger_c(g, Cc)
        return cls(sliding_law=SLIDING_LAWS[law], v0=v0, Cx=Cx, Cy=Cy, **kw)

    def c(self) -> _Model:
        if self.sliding_law >= 2 and (self.Cx is None or self.Cy is None):
            raise ValueError("the field sliding laws need the face coefficients Cx, Cy (Model.field_law)")
        nul = C.POINTER(C.c_double)()
        return _Model(self.kfs, self.gap_scheme, self.sliding_law, self.sliding_C, self.v0,
                      _p(self.Cx) if self.Cx is not None else nul, _p(self.Cy) if self.Cy is not None else nul)


class State:
    """Every field of src/state.jl:13-57 as Fortran-ordered float64 arrays."""

    def __init__(self, g: Grid):
        nx, ny = g.nx, g.ny
        for f in CENTER_FIELDS:
            setattr(self, f, np.zeros((nx, ny), order="F"))
        for f in XFACE_FIELDS:
            setattr(self, f, np.zeros((nx + 1, ny), order="F"))
        for f in YFACE_FIELDS:
            setattr(self, f, np.zeros((nx, ny + 1), order="F"))
        self.valid_x[:] = 1.0
        self.valid_y[:] = 1.0

    def c(self) -> _State:
        return _State(*[_p(getattr(self, f)) for f in ALL_FIELDS])

    def copy(self) -> "State":
        new = State.__new__(State)
        for f in ALL_FIELDS:
            setattr(new, f, np.asfortranarray(getattr(self, f).copy()))
        return new


def F(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def set_initial_conditions(s: State, g: Grid, p: ModelParameters, m: Model, mask, A_visc, zb, zs, b, G, ub_x, ub_y,
                           ieb, taub_x, taub_y) -> State:
    """src/initial_conditions.jl:14-126."""
    for name, val in [("mask", mask), ("A_visc", A_visc), ("zb", zb), ("zs", zs), ("b", b), ("G", G),
                      ("ub_x", ub_x), ("ub_y", ub_y), ("ieb", ieb), ("taub_x", taub_x), ("taub_y", taub_y)]:
        getattr(s, name)[...] = F(val)
    lib().orc_set_initial_conditions(C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()))
    return s


def update_derived(s: State, g: Grid, p: ModelParameters, m: Model) -> State:
    """The tail of Picard_iteration! (src/elliptic_solver.jl:193-206)."""
    lib().orc_update_derived(C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()))
    return s


def test_refresh_lagged(s: State, g: Grid, p: ModelParameters, m: Model) -> State:
    """The refresh sequence of test/linear_solver_test.jl:38-54 (lagged q/Re)."""
    lib().orc_test_refresh_lagged(C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()))
    return s


def step_b(s: State, g: Grid, p: ModelParameters, m: Model, dt: float) -> State:
    """src/run.jl:154-164."""
    lib().orc_step_b(C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()), dt)
    return s


def assemble_mfls(s: State, g: Grid, p: ModelParameters, kfs: int, parabolic: bool = False, dt: float = 0.0):
    """src/linear_solver.jl:438-610. Returns (aP, aE, aW, aN, aS, rhs) with rhs as an (nx, ny) array."""
    out = [np.zeros((g.nx, g.ny), order="F") for _ in range(6)]
    lib().orc_assemble_mfls(C.byref(g.c()), C.byref(p.c()), kfs, C.byref(s.c()), int(parabolic), dt,
                            *[_p(a) for a in out])
    return tuple(out)


def assemble_sals(s: State, g: Grid, p: ModelParameters, kfs: int, parabolic: bool = False, dt: float = 0.0):
    """src/linear_solver.jl:144-430 as a scipy CSR matrix plus rhs (flat, row = ix + iy*nx)."""
    import scipy.sparse as sp
    n = g.nx * g.ny
    cols = np.zeros(5 * n, dtype=np.int64)
    vals = np.zeros(5 * n)
    rhs = np.zeros(n)
    lib().orc_assemble_sals(C.byref(g.c()), C.byref(p.c()), kfs, C.byref(s.c()), int(parabolic), dt,
                            cols.ctypes.data_as(C.POINTER(C.c_int64)), vals.ctypes.data_as(C.POINTER(C.c_double)),
                            rhs.ctypes.data_as(C.POINTER(C.c_double)))
    rows = np.repeat(np.arange(n), 5)
    keep = cols >= 0
    M = sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(n, n))
    return M, rhs


def matvec(coeffs, x: np.ndarray) -> np.ndarray:
    aP = coeffs[0]
    nx, ny = aP.shape
    y = np.zeros((nx, ny), order="F")
    lib().orc_matvec(nx, ny, *[_p(a) for a in coeffs[:5]], _p(F(x).reshape((nx, ny), order="F")), _p(y))
    return y


def _op(coeffs, d=None) -> _Op:
    aP = coeffs[0]
    nx, ny = aP.shape
    return _Op(nx, ny, *[_p(a) for a in coeffs[:5]], _p(aP if d is None else d))


def tridiag_extreme_eigs(a, e):
    a = np.ascontiguousarray(a, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    lo, hi = C.c_double(), C.c_double()
    lib().orc_tridiag_extreme_eigs(len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
                                   e.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def estimate_eigenvalue_bounds(coeffs, rhs, nsteps=15, d=None):
    """src/preconditioner.jl:89-134."""
    lo, hi = C.c_double(), C.c_double()
    op = _op(coeffs, d)
    lib().orc_estimate_eigenvalue_bounds(C.byref(op), _p(F(rhs)), nsteps, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def estimate_eigenvalue_bounds_d(coeffs, rhs, nsteps=15, d=None):
    """The estimate in the D-inner product (Jacobi-PCG's own recurrence): 0.9 /
    1.1 x Ritz values that lie in the spectrum of D^-1 A."""
    lo, hi = C.c_double(), C.c_double()
    op = _op(coeffs, d)
    lib().orc_estimate_eigenvalue_bounds_d(C.byref(op), _p(F(rhs)), nsteps, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def update_chebyshev_bounds(coeffs, rhs, bounds=(1.0, 1.0), nsteps=15, d=None):
    """src/preconditioner.jl:197-233: returns (lambda_min, lambda_max, warned)."""
    lo, hi = C.c_double(bounds[0]), C.c_double(bounds[1])
    op = _op(coeffs, d)
    w = lib().orc_update_chebyshev_bounds(C.byref(op), _p(F(rhs)), nsteps, C.byref(lo), C.byref(hi))
    return lo.value, hi.value, bool(w)


def cheb_apply(coeffs, x, degree, lmin, lmax, d=None):
    """src/preconditioner.jl:239-267."""
    nx, ny = coeffs[0].shape
    y, r, pp, Ap = [np.zeros((nx, ny), order="F") for _ in range(4)]
    op = _op(coeffs, d)
    lib().orc_cheb_apply(C.byref(op), degree, lmin, lmax, _p(F(x).reshape((nx, ny), order="F")), _p(y), _p(r),
                         _p(pp), _p(Ap))
    return y


def gershgorin(coeffs, d=None) -> float:
    """1 + max_i sum_j |a_ij| / |d_i|, the Gershgorin bound of D^-1 A."""
    op = _op(coeffs, d)
    return lib().orc_gershgorin(C.byref(op))


def cg_solve(coeffs, rhs, x0, precond="jacobi", degree=4, nsteps=15, atol=1.4901161193847656e-8,
             rtol=1.4901161193847656e-8, itmax=0, bounds=(1.0, 1.0), gershgorin=False):
    """Krylov-semantics PCG as called at src/linear_solver.jl:915-945. Returns (x, stats, bounds).
    gershgorin=True takes lambda_max from the Gershgorin bound (the HIP library's default bounds)."""
    nx, ny = coeffs[0].shape
    cfg = _CgCfg(1 if precond == "chebyshev" else 0, degree, nsteps, itmax, atol, rtol, bounds[0], bounds[1],
                 1 if gershgorin else 0, 0)
    x = F(x0).reshape((nx, ny), order="F").copy(order="F")
    st = CgStats()
    op = _op(coeffs)
    lib().orc_cg_solve(C.byref(op), C.byref(cfg), _p(F(rhs).reshape((nx, ny), order="F")), _p(x), C.byref(st))
    return x, st, (cfg.lambda_min, cfg.lambda_max)


def cholesky_solve(coeffs, rhs):
    """Band-Cholesky stand-in for CHOLMOD (src/linear_solver.jl:709-739)."""
    nx, ny = coeffs[0].shape
    x = np.zeros((nx, ny), order="F")
    info = lib().orc_cholesky_solve(nx, ny, *[_p(a) for a in coeffs[:5]],
                                    _p(F(rhs).reshape((nx, ny), order="F")), _p(x))
    if info != 0:
        raise np.linalg.LinAlgError(f"band Cholesky failed at pivot {info}")
    return x


class Solver:
    """CholeskyDirectSolver (kind='cholesky') or CGIterativeSolver over the
    matrix-free system (kind='cg'), src/linear_solver.jl:131-221, 760-843."""

    def __init__(self, g: Grid, kind: str = "cholesky", precond: str = "jacobi", degree: int = 4,
                 nsteps: int = 15, atol: float = 1.4901161193847656e-8, rtol: float = 1.4901161193847656e-8,
                 itmax: int = 0, gershgorin: bool = False):
        if precond == "chebyshev" and degree < 1:
            raise ValueError("ChebyshevPreconditioner: degree must be >= 1")
        if precond == "chebyshev" and nsteps < 2:
            raise ValueError("ChebyshevPreconditioner: nsteps_estimate must be >= 2")
        cfg = _CgCfg(1 if precond == "chebyshev" else 0, degree, nsteps, itmax, atol, rtol, 1.0, 1.0,
                     1 if gershgorin else 0, 0)
        self.kind = kind
        self.g = g
        self._h = lib().orc_solver_create(g.nx, g.ny, 0 if kind == "cholesky" else 1, C.byref(cfg))
        self.last_stats = CgStats()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_solver_destroy(self._h)
            self._h = None

    @property
    def bounds(self):
        lo, hi = C.c_double(), C.c_double()
        lib().orc_solver_bounds(self._h, C.byref(lo), C.byref(hi))
        return lo.value, hi.value

    def solve_elliptic(self, s: State, p: ModelParameters, kfs: int) -> State:
        rc = lib().orc_solve_linear(self._h, C.byref(self.g.c()), C.byref(p.c()), kfs, C.byref(s.c()), 0, 0.0,
                                    C.byref(self.last_stats))
        if rc != 0:
            raise np.linalg.LinAlgError("oracle linear solve failed")
        return s

    def solve_parabolic(self, s: State, p: ModelParameters, kfs: int, dt: float) -> State:
        rc = lib().orc_solve_linear(self._h, C.byref(self.g.c()), C.byref(p.c()), kfs, C.byref(s.c()), 1, dt,
                                    C.byref(self.last_stats))
        if rc != 0:
            raise np.linalg.LinAlgError("oracle linear solve failed")
        return s


def picard_loop(sv: Solver, s: State, g: Grid, p: ModelParameters, m: Model, iters=500, tol=1e-6, check_every=1,
                alpha=None):
    """src/elliptic_solver.jl:138-176. Returns (converged, last_iter, cg_iters)."""
    last = C.c_int32()
    cgi = C.c_int64(0)
    conv = lib().orc_picard_loop(sv._h, C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()), iters, tol,
                                 check_every, -1.0 if alpha is None else float(alpha), C.byref(last), C.byref(cgi))
    if conv < 0:
        raise np.linalg.LinAlgError("oracle linear solve failed (CG breakdown or failed factorisation)")
    return bool(conv), last.value, cgi.value


def parabolic_step(sv: Solver, s: State, g: Grid, p: ModelParameters, m: Model, dt: float) -> int:
    """src/parabolic_solver.jl:16-38. Returns CG iterations used."""
    cgi = C.c_int64(0)
    rc = lib().orc_parabolic_step(sv._h, C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()), dt,
                                  C.byref(cgi))
    if rc != 0:
        raise np.linalg.LinAlgError("oracle linear solve failed (CG breakdown or failed factorisation)")
    return cgi.value


def seasonal_ieb(t, t_start=0.4, t_end=0.7, amplitude=492.75, offset=493.75, i_min=1.0,
                 seconds_per_year=365 * 86400.0):
    """src/melt_input.jl:66-94: uniform ieb value at elapsed time t."""
    return lib().orc_seasonal_ieb(t, t_start, t_end - t_start, amplitude, offset, i_min, seconds_per_year)


def step(sv: Solver, s: State, g: Grid, p: ModelParameters, m: Model, dt: float, ieb_mode=0, ieb_uniform=0.0,
         moulin=None, iters=500, tol=1e-6, check_every=1, alpha=None):
    """src/run.jl:82-88. Returns (converged, last_iter, cg_iters)."""
    last = C.c_int32()
    cgi = C.c_int64(0)
    mp = _p(F(moulin)) if moulin is not None else C.cast(0, C.POINTER(C.c_double))
    conv = lib().orc_step(sv._h, C.byref(g.c()), C.byref(p.c()), C.byref(m.c()), C.byref(s.c()), dt, ieb_mode,
                          ieb_uniform, mp, iters, tol, check_every, -1.0 if alpha is None else float(alpha),
                          C.byref(last), C.byref(cgi))
    if conv < 0:
        raise np.linalg.LinAlgError("oracle linear solve failed (CG breakdown or failed factorisation)")
    return bool(conv), last.value, cgi.value


def num_threads() -> int:
    return lib().orc_num_threads()


def test_fixture_6x6(sliding_C=0.25, kfs=0, e_v=0.0, b_min=0.0):
    """The shared 6x6 mixed-mask case of the reference tests
    (test/linear_solver_test.jl:8-32, test/parabolic_solver_test.jl:4-23)."""
    nx, ny = 6, 6
    g = Grid(nx, ny, 1e3, 1e3)
    p = ModelParameters(e_v=e_v, b_min=b_min)
    m = Model(kfs=kfs, sliding_C=sliding_C)
    mask = np.full((nx, ny), GROUNDED)
    mask[-1, :] = OCEAN
    mask[0, :] = LAND
    mask[:, 0] = OTHER_BASIN
    A_visc = np.full((nx, ny), 5e-25)
    zb = np.repeat((-0.02 * g.x).reshape(nx, 1), ny, axis=1)
    zs = zb + 500.0
    b = np.full((nx, ny), 0.01)
    G = np.full((nx, ny), 0.06)
    ub_x = np.full((nx + 1, ny), 1e-6)
    ub_y = np.zeros((nx, ny + 1))
    ieb = np.zeros((nx, ny))
    ieb[2, 2] = 3 / (g.dx * g.dy)
    taub_x = np.zeros((nx + 1, ny))
    taub_y = np.zeros((nx, ny + 1))
    raw = dict(mask=mask, A_visc=A_visc, zb=zb, zs=zs, b=b, G=G, ub_x=ub_x, ub_y=ub_y, ieb=ieb, taub_x=taub_x,
               taub_y=taub_y)
    return g, p, m, raw


def perturb_like_reference_tests(s: State, g: Grid, p: ModelParameters, m: Model) -> State:
    """state.h .+= 0.01 .* reshape(1:(nx*ny), nx, ny) then the lagged refresh
    (test/linear_solver_test.jl:38-54)."""
    s.h += 0.01 * np.arange(1, g.nx * g.ny + 1, dtype=np.float64).reshape((g.nx, g.ny), order="F")
    return test_refresh_lagged(s, g, p, m)


if __name__ == "__main__":
    build(force=True)
    print("oracle built:", _LIB_PATH, "threads:", num_threads(), os.environ.get("OMP_NUM_THREADS"))


RAW_INPUTS = ["mask", "A_visc", "zb", "zs", "b", "G", "ub_x", "ub_y", "ieb", "taub_x", "taub_y"]


def generate_glacier(g: Grid, seed: int = 0, bed_amp: float = 10.0, corr_len: float = 2000.0,
                     trough_width: float = 0.0, trough_depth: float = 0.0, b0: float = 0.005, A: float = 2.5e-25,
                     ub: float = 1e-6, G: float = 0.05, ieb: float = 1e-8, n_moulins: int = 0, qmin: float = 1.0,
                     qmax: float = 10.0):
    """The device glacier generator restated (orc_generate_glacier): returns
    (raw inputs dict for set_initial_conditions, moulin field or None)."""
    s = State(g)
    moulin = np.zeros((g.nx, g.ny), order="F") if n_moulins > 0 else None
    spec = _GlacierSpec(int(seed), bed_amp, corr_len, trough_width, trough_depth, b0, A, ub, G, ieb,
                        int(n_moulins), 0, qmin, qmax)
    rc = lib().orc_generate_glacier(C.byref(g.c()), C.byref(spec), C.byref(s.c()),
                                    _p(moulin) if moulin is not None else C.cast(0, C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError("bad generator arguments")
    return {n: getattr(s, n) for n in RAW_INPUTS}, moulin


def splitmix64(z: int) -> int:
    return int(lib().orc_splitmix64(z))
