"""Host-side mirror of Shakti.jl's hydraulic-head interface over libshakti_hip.so.

Names, argument meaning and error behaviour follow the reference (Julia's
trailing ``!`` is dropped): ``State(g)`` allocates a device-resident State,
``set_initial_conditions`` fills it, ``CGIterativeSolver(g,
HIPMatrixFreeLinearSystem; chebyshev_degree=...)`` is the new linear-solver
subtype that ``PicardSolver`` / ``Simulation`` / ``run`` drive unchanged.
Every computation runs in the HIP library; this module only moves arguments.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
import weakref
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L
from . import observer as O

GROUNDED, OCEAN, LAND, OTHER_BASIN = 0.0, 1.0, 2.0, 3.0  # src/mask.jl:21-41
SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)            # Krylov's default atol/rtol


# ------------------------------------------------------------------ data model

class Grid:
    """src/grid.jl:26-35: nx*ny cells over lx*ly, dx = lx/(nx-1)."""

    def __init__(self, nx: int, ny: int, lx: float, ly: float):
        self.nx, self.ny = int(nx), int(ny)
        self.lx, self.ly = float(lx), float(ly)
        self.dx = self.lx / (self.nx - 1)
        self.dy = self.ly / (self.ny - 1)
        self.x = np.arange(self.nx) * self.dx
        self.y = np.arange(self.ny) * self.dy
        self.dx2 = self.dx ** 2
        self.dy2 = self.dy ** 2

    def c(self) -> L.ShkGrid:
        return L.ShkGrid(self.nx, self.ny, self.lx, self.ly, self.dx, self.dy, self.dx2, self.dy2)


@dataclass
class ModelParameters:
    """src/model_parameters.jl:55-84 (keyword constructor defaults)."""
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

    def c(self) -> L.ShkParams:
        return L.ShkParams(*[float(getattr(self, f.name)) for f in fields(self)])


class AbstractKFaceScheme:
    code = -1


class Arithmetic(AbstractKFaceScheme):
    """src/k_face_scheme.jl:15."""
    code = 0


class Harmonic(AbstractKFaceScheme):
    """src/k_face_scheme.jl:23."""
    code = 1


class AbstractSlidingLaw:
    pass


@dataclass
class RegularizedCoulombSlidingLaw(AbstractSlidingLaw):
    """src/sliding_law.jl:33-35."""
    C: float


class PrescribedSlidingLaw(AbstractSlidingLaw):
    """src/sliding_law.jl:45: taub fixed at set-up."""


class _FieldSlidingLaw(AbstractSlidingLaw):
    """A law whose coefficient C is a field given at the cell centres (an (nx, ny) array) or a
    uniform scalar.  The constructors of the reference stagger C onto the faces on the host; here
    the cell values are kept and staggered on the device (shk_set_sliding_coefficient) when a State
    first uses the law, with the same face values (outer faces the cell's value, inner faces the
    mean of the two cells).  Like the reference's immutable laws, a law object's C is fixed once a
    State has used it: build a new law to change C."""
    code = -1

    def __init__(self, g: Grid, C):
        a = np.asarray(C, dtype=np.float64)
        if a.ndim == 0:
            a = np.full((g.nx, g.ny), float(a))
        if a.shape != (g.nx, g.ny):
            raise ValueError(f"C must be a scalar or an array of shape {(g.nx, g.ny)}, not {a.shape}")
        self.grid = g
        self.C_cells = np.asfortranarray(a)


class LinearSlidingLaw(_FieldSlidingLaw):
    """LinearSlidingLaw(g, C) (src/sliding_law.jl:61-90): taub = C^2 N ub on every face."""
    code = 2


class RegularizedCoulombFieldSlidingLaw(_FieldSlidingLaw):
    """RegularizedCoulombFieldSlidingLaw(g, C) (src/sliding_law.jl:270-301): the regularised-Coulomb
    stress with a field C."""
    code = 3


class RegularizedCoulombV0SlidingLaw(_FieldSlidingLaw):
    """RegularizedCoulombV0SlidingLaw(g, C, v0) (src/sliding_law.jl:422-452): taub =
    N C (|ub| / (|ub| + v0))^(1/n) ub / |ub|, v0 in the units of ub (m/s)."""
    code = 4

    def __init__(self, g: Grid, C, v0: float):
        super().__init__(g, C)
        self.v0 = float(v0)


class AbstractMeltInput:
    def value(self, t: float):
        return None


class ConstantMeltInput(AbstractMeltInput):
    """src/melt_input.jl:22,30: ieb never changes."""


@dataclass
class SeasonalMeltInput(AbstractMeltInput):
    """src/melt_input.jl:49-94: uniform ieb following a cosine melt season."""
    t_start: float = 0.4
    t_end: float = 0.7
    amplitude: float = 492.75
    offset: float = 493.75
    i_min: float = 1.0
    seconds_per_year: float = 365 * 86400.0

    def value(self, t: float) -> float:
        period = self.t_end - self.t_start
        omega = 2 * math.pi / period
        yf = math.fmod(t / self.seconds_per_year, 1.0)
        if yf < 0:
            yf += 1.0
        if self.t_start <= yf <= self.t_start + period:
            i_ma = self.offset - self.amplitude * math.cos(omega * (yf - self.t_start))
        else:
            i_ma = self.i_min
        return i_ma / self.seconds_per_year


class SeasonalMoulinMeltInput(SeasonalMeltInput):
    """A SeasonalMeltInput plus a static field of point moulins added on top
    (ieb = uniform(t) + moulin).  No counterpart in the reference, whose
    update_ieb! overwrites ieb uniformly (src/melt_input.jl:92).  moulin None
    uses the field already on the device (synthetic.generate_on_device)."""

    def __init__(self, moulin: np.ndarray | None, **kw):
        super().__init__(**kw)
        self.moulin = None if moulin is None else np.asfortranarray(np.asarray(moulin, dtype=np.float64))


# ------------------------------------------------------------------ State

def _finalize_ctx(handle):
    if handle.value:
        L.load().shk_destroy(handle)
        handle.value = None


class State:
    """A device-resident State (src/state.jl:13-137).  Field reads download a
    Fortran-ordered numpy copy (``state.h[ix, iy]``); ``state.set(name, a)``
    uploads.  The context created here also holds the solver workspace."""

    def __init__(self, g: Grid, p: ModelParameters | None = None):
        self.grid = g
        lib = L.load()
        self._ctx = C.c_void_p()
        cfg = _solver_cfg(None)
        m = L.ShkModel(0, 0, 0, 0.5, 0.0)
        L.check(lib.shk_create(C.byref(g.c()), C.byref((p or ModelParameters()).c()), C.byref(m), C.byref(cfg),
                               C.byref(self._ctx)), None, "shk_create")
        weakref.finalize(self, _finalize_ctx, self._ctx)

    @property
    def ctx(self):
        return self._ctx

    def shape_of(self, name: str):
        i = L.FIELD_ID[name]
        nx, ny = self.grid.nx, self.grid.ny
        if i < L.N_CENTER:
            return (nx, ny)
        if i < L.N_CENTER + L.N_XFACE:
            return (nx + 1, ny)
        return (nx, ny + 1)

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape_of(name), order="F")
        L.check(L.load().shk_download(self._ctx, L.FIELD_ID[name], out.ctypes.data_as(C.POINTER(C.c_double)),
                                      out.size), self._ctx, f"download {name}")
        return out

    def set(self, name: str, value) -> None:
        shape = self.shape_of(name)
        a = np.asfortranarray(np.broadcast_to(np.asarray(value, dtype=np.float64), shape))
        L.check(L.load().shk_upload(self._ctx, L.FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), a.size),
                self._ctx, f"upload {name}")

    def __getattr__(self, name):
        if name in L.FIELD_ID:
            return self.get(name)
        raise AttributeError(name)

    def to_dict(self) -> dict:
        return {n: self.get(n) for n in L.FIELD_NAMES}

    def snapshot(self, names=None) -> None:
        """Copy the named fields (every field for None) on the device and queue
        their copy to pinned host memory (shk_snapshot); the State may change
        once this returns.  ``fetch`` hands them out."""
        if names is None:
            L.check(L.load().shk_snapshot(self._ctx, None, 0), self._ctx, "shk_snapshot")
            return
        ids = (C.c_int32 * len(names))(*[L.FIELD_ID[n] for n in names])
        L.check(L.load().shk_snapshot(self._ctx, ids, len(names)), self._ctx, "shk_snapshot")

    def fetch(self, name: str) -> np.ndarray:
        """One field of the last snapshot (waits for its copy)."""
        out = np.empty(self.shape_of(name), order="F")
        L.check(L.load().shk_snapshot_fetch(self._ctx, L.FIELD_ID[name], out.ctypes.data_as(C.POINTER(C.c_double)),
                                            out.size), self._ctx, f"shk_snapshot_fetch {name}")
        return out

    def stats(self, name: str) -> tuple[float, float, float]:
        """(min, max, mean) of a field, reduced on the device (shk_field_stats)."""
        out = (C.c_double * 3)()
        L.check(L.load().shk_field_stats(self._ctx, L.FIELD_ID[name], out), self._ctx, f"shk_field_stats {name}")
        return float(out[0]), float(out[1]), float(out[2])

    # per-kernel HIP-event timing inside the library (see shk_profile_*)
    PROFILE_IDS = ["assemble", "lanczos", "cheb_bounds", "cg_init", "cheb_pass", "cg_matvec", "cg_update",
                   "cg_finish", "derived", "gap", "stagger_b", "ieb", "ic", "matvec", "ir", "cg_update_cheb",
                   "mg_down", "mg_up", "mg_tail", "mg_setup", "nn_forward", "nn_warm_start", "mg_down_fine",
                   "mg_up_fine", "nn_precond"]

    def profile_enable(self, kernels=None) -> None:
        """kernels: None for all, [] to switch off, or a list of kernel names."""
        if kernels is None:
            mask = 1
        else:
            mask = 0
            for k in kernels:
                mask |= 1 << self.PROFILE_IDS.index(k)
            if mask == 1:  # kernel 0 alone would read as "all"; widen to a bit that is never used
                mask |= 1 << 30
        L.check(L.load().shk_profile_enable(self._ctx, mask), self._ctx, "shk_profile_enable")

    def profile_read(self) -> dict:
        n = 32
        names = (C.c_char_p * n)()
        counts = (C.c_int64 * n)()
        ms = (C.c_double * n)()
        by = (C.c_double * n)()
        k = C.c_int32()
        L.check(L.load().shk_profile_read(self._ctx, n, names, counts, ms, by, C.byref(k)), self._ctx,
                "shk_profile_read")
        return {names[i].decode(): {"count": counts[i], "ms": ms[i], "bytes": by[i]} for i in range(k.value)
                if counts[i] > 0}

    def profile_reset(self) -> None:
        L.check(L.load().shk_profile_reset(self._ctx), self._ctx, "shk_profile_reset")

    def download_system(self) -> dict:
        """The last assembly in the reference's five-array layout
        (shk_download_system): aP, aE, aW, aN, aS and rhs, each (nx, ny)."""
        shape = (self.grid.nx, self.grid.ny)
        arrs = [np.zeros(shape, order="F") for _ in range(6)]
        L.check(L.load().shk_download_system(self._ctx, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]),
                self._ctx, "shk_download_system")
        return dict(zip(["aP", "aE", "aW", "aN", "aS", "rhs"], arrs))

    def solve_log(self, max_entries: int = 65536) -> list[dict]:
        """Statistics of every linear solve since the last call, oldest first
        (shk_solve_log): those inside the Picard loop and whole steps too."""
        buf = (L.ShkSolveStats * max_entries)()
        k = C.c_int32()
        L.check(L.load().shk_solve_log(self._ctx, buf, max_entries, C.byref(k)), self._ctx, "shk_solve_log")
        return [buf[i].as_dict() for i in range(k.value)]

    # internal: push the per-call choices before an entry point
    def _bind(self, p: ModelParameters | None = None, kfs: AbstractKFaceScheme | None = None,
              sl: AbstractSlidingLaw | None = None, gap_scheme: int | None = None):
        lib = L.load()
        if p is not None:
            L.check(lib.shk_set_params(self._ctx, C.byref(p.c())), self._ctx, "shk_set_params")
        if kfs is not None or sl is not None or gap_scheme is not None:
            m = getattr(self, "_model", L.ShkModel(0, 0, 0, 0.5, 0.0))
            if kfs is not None:
                m.kfs = kfs.code
            if sl is not None:
                if isinstance(sl, RegularizedCoulombSlidingLaw):
                    m.sliding_law, m.sliding_C = 0, float(sl.C)
                elif isinstance(sl, PrescribedSlidingLaw):
                    m.sliding_law = 1
                elif isinstance(sl, _FieldSlidingLaw):
                    if (sl.grid.nx, sl.grid.ny) != (self.grid.nx, self.grid.ny):
                        raise ValueError("the sliding law was built on another grid")
                    m.sliding_law, m.sliding_v0 = sl.code, getattr(sl, "v0", 0.0)
                    # the coefficient goes up once per law object
                    if getattr(self, "_sliding_bound", None) is not sl:
                        c = sl.C_cells
                        L.check(lib.shk_set_sliding_coefficient(self._ctx, c.ctypes.data_as(C.POINTER(C.c_double)),
                                                                c.size), self._ctx, "shk_set_sliding_coefficient")
                        self._sliding_bound = sl
                else:
                    raise L.ShaktiHIPError(f"unsupported sliding law {type(sl).__name__}")
            if gap_scheme is not None:
                m.gap_scheme = gap_scheme
            self._model = m
            L.check(lib.shk_set_model(self._ctx, C.byref(m)), self._ctx, "shk_set_model")


def set_initial_conditions(s: State, g: Grid, p: ModelParameters, sl: AbstractSlidingLaw, mask, A_visc, zb, zs, b, G,
                           ub_x, ub_y, ieb, taub_x, taub_y) -> State:
    """set_initial_conditions! (src/initial_conditions.jl:14-126), on the device."""
    s._bind(p=p, sl=sl)
    for name, val in [("mask", mask), ("A_visc", A_visc), ("zb", zb), ("zs", zs), ("b", b), ("G", G),
                      ("ub_x", ub_x), ("ub_y", ub_y), ("ieb", ieb), ("taub_x", taub_x), ("taub_y", taub_y)]:
        s.set(name, val)
    L.check(L.load().shk_set_initial_conditions(s.ctx), s.ctx, "shk_set_initial_conditions")
    return s


def set_initial_conditions_on_device(s: State, g: Grid, p: ModelParameters, sl: AbstractSlidingLaw) -> State:
    """set_initial_conditions! from the raw inputs already in the device
    State (written by synthetic.generate_on_device or uploaded field by
    field): nothing crosses PCIe."""
    s._bind(p=p, sl=sl)
    L.check(L.load().shk_set_initial_conditions(s.ctx), s.ctx, "shk_set_initial_conditions")
    return s


# ------------------------------------------------------------- linear solver

class AbstractLinearSystem:
    pass


class HIPMatrixFreeLinearSystem(AbstractLinearSystem):
    """Device-resident matrix-free system in the symmetric face layout
    (aP, aE, aN, rhs); the counterpart of MatrixFreeLinearSystem
    (src/linear_solver.jl:91-121)."""


class MatrixFreeLinearSystem(HIPMatrixFreeLinearSystem):
    """Alias so reference scripts that name MatrixFreeLinearSystem run here."""


class SparseAssembledLinearSystem(AbstractLinearSystem):
    """CPU-only in the reference; refused here as it is refused on GPU backends."""


@dataclass
class _Stats:
    """Krylov's ws.stats of the last solve, plus what the library measures
    beside it (breakdown, bound estimate, FP64 residual of the result)."""
    niter: int = 0
    solved: bool = False
    zero_curvature: bool = False
    rnorm0: float = 0.0
    rnorm: float = 0.0
    refinements: int = 0
    fallback: int = 0  # bit 0: warm start rejected; bit 1: learned preconditioner stalled
    tired: bool = False
    true_residual: float = float("nan")  # ||b - A h|| / ||b|| of the returned head (FP64)


class _Workspace:
    def __init__(self):
        self.stats = _Stats()


@dataclass
class ChebyshevPreconditioner:
    """The bounds the last solve used (src/preconditioner.jl:160-186)."""
    degree: int
    nsteps_estimate: int = 15
    lambda_min: float = 1.0
    lambda_max: float = 1.0


class AbstractLinearSolver:
    pass


class AbstractIterativeSolver(AbstractLinearSolver):
    pass


FORMS = {"fp64": 0, "fp32_refine": 1, "learned": 2}


class CGIterativeSolver(AbstractIterativeSolver):
    """CGIterativeSolver(g, HIPMatrixFreeLinearSystem; chebyshev_degree,
    chebyshev_nsteps_estimate) (src/linear_solver.jl:832-843) on the GPU.
    ``atol``/``rtol``/``itmax`` are Krylov's cg! stopping controls (defaults
    sqrt(eps), sqrt(eps), 2n); ``check_every`` is how many CG iterations are
    queued between host looks at the device-side convergence flag.

    ``multigrid=True`` preconditions with the FP32 geometric-multigrid
    V-cycle: with ``form="fp64"`` it preconditions FP64 CG (residuals, dots
    and updates stay FP64); with ``form="fp32_refine"`` it preconditions the
    FP32 inner solves of the FP64 refinement loop.

    ``form="learned"`` with ``network=`` a trained ``learned.UNet`` runs FP64
    CG (with the chosen preconditioner) from the network's warm start; the
    stopping test still refers to the residual at the Picard iterate h, and
    a warm start that is not a descent step is dropped (``ws.stats.fallback``
    bit 0).  ``learned_precond_every=m`` (m >= 2) also makes the network a
    preconditioner inside flexible CG: one iteration in m preconditions with
    the network's correction composed with the fixed preconditioner; a
    batch of m iterations that does not lower the residual norm below
    ``stall_ratio`` times its start value ends that, and the solve goes on as
    plain PCG (``ws.stats.fallback`` bit 1)."""

    def __init__(self, g: Grid, system=HIPMatrixFreeLinearSystem, chebyshev_degree: int | None = None,
                 chebyshev_nsteps_estimate: int = 15, amg: bool = False, atol: float = SQRT_EPS,
                 rtol: float = SQRT_EPS, itmax: int = 0, form: str = "fp64", check_every: int = 8,
                 inner_rtol: float = 1e-3, multigrid: bool = False, network=None,
                 learned_precond_every: int = 0, stall_ratio: float = 1.0, chebyshev_bounds: str = "gershgorin",
                 picard_atol: bool = False):
        if system is SparseAssembledLinearSystem:
            raise L.ShaktiHIPError("CGIterativeSolver(g, SparseAssembledLinearSystem) is CPU-only; use "
                                   "HIPMatrixFreeLinearSystem on the GPU.")
        if not (isinstance(system, type) and issubclass(system, HIPMatrixFreeLinearSystem)):
            raise L.ShaktiHIPError(f"unsupported linear system {system}")
        if amg:
            raise L.ShaktiHIPError("AMGPreconditioner is CPU/SparseMatrixCSC-only; use chebyshev_degree here.")
        if chebyshev_degree is not None and chebyshev_degree < 1:
            raise L.ShaktiHIPError(f"ChebyshevPreconditioner: degree must be >= 1 (got {chebyshev_degree})")
        if chebyshev_nsteps_estimate < 2:
            raise L.ShaktiHIPError(
                f"ChebyshevPreconditioner: nsteps_estimate must be >= 2 (got {chebyshev_nsteps_estimate})")
        if form not in FORMS:
            raise L.ShaktiHIPError(f"unknown solver form {form!r}")
        if chebyshev_bounds not in L.CHEB_BOUNDS:
            raise L.ShaktiHIPError(f"chebyshev_bounds must be 'gershgorin' or 'reference' (got {chebyshev_bounds!r})")
        if multigrid and chebyshev_degree is not None:
            raise L.ShaktiHIPError("multigrid=True replaces chebyshev_degree: give one of them")
        if (form == "learned") != (network is not None):
            raise L.ShaktiHIPError("form='learned' takes the trained network (shakti_hip.learned.UNet) as "
                                   "network=, and only that form does")
        if learned_precond_every and form != "learned":
            raise L.ShaktiHIPError("learned_precond_every needs form='learned' and its network")
        if learned_precond_every and not multigrid and chebyshev_degree is None:
            raise L.ShaktiHIPError("the learned preconditioner composes with multigrid=True or chebyshev_degree")
        self.network = network
        self.learned_precond_every = int(learned_precond_every)
        self.stall_ratio = float(stall_ratio)
        self.multigrid = bool(multigrid)
        self.grid = g
        self.chebyshev_degree = chebyshev_degree
        self.nsteps_estimate = chebyshev_nsteps_estimate
        self.atol, self.rtol, self.itmax = float(atol), float(rtol), int(itmax)
        self.form = form
        self.check_every = int(check_every)
        self.inner_rtol = float(inner_rtol)
        self.chebyshev_bounds = chebyshev_bounds
        # inside a Picard loop, later solves stop at rtol x the first solve's initial residual (SHK_CFG_PICARD_ATOL)
        self.picard_atol = bool(picard_atol)
        self.ws = _Workspace()
        self.precond = None if chebyshev_degree is None else ChebyshevPreconditioner(chebyshev_degree,
                                                                                     chebyshev_nsteps_estimate)
        self.last_stats = L.ShkSolveStats()

    def cfg(self) -> L.ShkSolverCfg:
        return _solver_cfg(self)

    def _record(self, st: L.ShkSolveStats):
        self.last_stats = st
        self.ws.stats = _Stats(st.niter, bool(st.solved), bool(st.zero_curvature), st.rnorm0, st.rnorm,
                               st.refinements, int(st.fallback), bool(st.tired), st.true_residual)
        if self.precond is not None:
            self.precond.lambda_min, self.precond.lambda_max = st.lambda_min, st.lambda_max
            if st.bounds_warning:
                import warnings
                warnings.warn("ChebyshevPreconditioner: eigenvalue bound estimate was invalid; reusing previous "
                              "bounds for this solve")


def _solver_cfg(ls: CGIterativeSolver | None) -> L.ShkSolverCfg:
    if ls is None:
        return L.ShkSolverCfg(0, 4, 15, 0, 0, 8, SQRT_EPS, SQRT_EPS, 1e-3, 0, 0)
    cheb = ls.chebyshev_degree is not None
    precond = 2 if ls.multigrid else (1 if cheb else 0)
    return L.ShkSolverCfg(precond, ls.chebyshev_degree or 1, ls.nsteps_estimate, FORMS[ls.form], ls.itmax,
                          ls.check_every, ls.atol, ls.rtol, ls.inner_rtol, L.CHEB_BOUNDS[ls.chebyshev_bounds],
                          1 if ls.picard_atol else 0)


def _bind_solver(s: State, ls: CGIterativeSolver):
    L.check(L.load().shk_set_solver_cfg(s.ctx, C.byref(ls.cfg())), s.ctx, "shk_set_solver_cfg")
    net = getattr(ls, "network", None)
    handle = net.handle if net is not None else None
    if getattr(s, "_attached", None) != handle:
        L.check(L.load().shk_attach_unet(s.ctx, handle), s.ctx, "shk_attach_unet")
        s._attached = handle
        s._attached_net = net  # the context holds the handle: keep the network alive
    L.check(L.load().shk_set_learned_precond(s.ctx, getattr(ls, "learned_precond_every", 0),
                                             getattr(ls, "stall_ratio", 1.0)), s.ctx, "shk_set_learned_precond")


def solve_elliptic_linear_system(ls: CGIterativeSolver, s: State, g: Grid, p: ModelParameters,
                                 kfs: AbstractKFaceScheme) -> None:
    """src/linear_solver.jl:915-925: assemble, precondition, PCG warm-started from h; s.h updated."""
    s._bind(p=p, kfs=kfs)
    _bind_solver(s, ls)
    st = L.ShkSolveStats()
    L.check(L.load().shk_solve_elliptic(s.ctx, C.byref(st)), s.ctx, "shk_solve_elliptic")
    ls._record(st)


def solve_parabolic_linear_system(ls: CGIterativeSolver, s: State, g: Grid, p: ModelParameters,
                                  kfs: AbstractKFaceScheme, dt: float) -> None:
    """src/linear_solver.jl:935-945."""
    s._bind(p=p, kfs=kfs)
    _bind_solver(s, ls)
    st = L.ShkSolveStats()
    L.check(L.load().shk_solve_parabolic(s.ctx, float(dt), C.byref(st)), s.ctx, "shk_solve_parabolic")
    ls._record(st)


def update_chebyshev_bounds(ls: CGIterativeSolver, s: State, rhs) -> ChebyshevPreconditioner:
    """update_chebyshev_bounds!(P, rhs) (src/preconditioner.jl:197-233) on the
    operator of the State's last assembly: a valid estimate from ``rhs``
    replaces ``ls.precond``'s bounds, an invalid one keeps them (the first time
    the placeholder (1, 1)) with a warning, as the reference does."""
    if ls.precond is None:
        raise L.ShaktiHIPError("update_chebyshev_bounds needs a CGIterativeSolver with chebyshev_degree")
    _bind_solver(s, ls)
    r = np.asfortranarray(np.broadcast_to(np.asarray(rhs, dtype=np.float64), (s.grid.nx, s.grid.ny)))
    lo, hi, w = C.c_double(), C.c_double(), C.c_int32()
    L.check(L.load().shk_update_chebyshev_bounds(s.ctx, r.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lo),
                                                 C.byref(hi), C.byref(w)), s.ctx, "shk_update_chebyshev_bounds")
    ls.precond.lambda_min, ls.precond.lambda_max = lo.value, hi.value
    if w.value:
        import warnings
        warnings.warn("ChebyshevPreconditioner: eigenvalue bound estimate was invalid; reusing previous bounds")
    return ls.precond


def update_derived(s: State, p: ModelParameters, sl: AbstractSlidingLaw) -> None:
    """The refresh after a solve (src/elliptic_solver.jl:193-206)."""
    s._bind(p=p, sl=sl)
    L.check(L.load().shk_update_derived(s.ctx), s.ctx, "shk_update_derived")


# ------------------------------------------------------------------ Picard

class PicardSolver:
    """src/elliptic_solver.jl:59-105."""

    def __init__(self, iters: int, tol: float, ls: AbstractLinearSolver, g: Grid, alpha: float | None = None,
                 check_every: int = 1):
        if not isinstance(ls, CGIterativeSolver):
            raise L.ShaktiHIPError("PicardSolver here needs a CGIterativeSolver over HIPMatrixFreeLinearSystem")
        self.iters, self.tol, self.ls = int(iters), float(tol), ls
        self.alpha = alpha
        self.check_every = int(check_every)
        self.converged = False
        self.last_iter = 0
        self.cg_iters = 0


def Picard_loop(ps: PicardSolver, s: State, g: Grid, p: ModelParameters, kfs: AbstractKFaceScheme,
                sl: AbstractSlidingLaw) -> None:
    """src/elliptic_solver.jl:138-176 (each iteration: :187-208)."""
    s._bind(p=p, kfs=kfs, sl=sl)
    _bind_solver(s, ps.ls)
    conv, last, cgi = C.c_int32(), C.c_int32(), C.c_int64(0)
    L.check(L.load().shk_picard_loop(s.ctx, ps.iters, ps.tol, ps.check_every,
                                     -1.0 if ps.alpha is None else float(ps.alpha), C.byref(conv), C.byref(last),
                                     C.byref(cgi)), s.ctx, "shk_picard_loop")
    ps.converged, ps.last_iter, ps.cg_iters = bool(conv.value), last.value, cgi.value


def elliptic_solver(ps: PicardSolver, s: State, g: Grid, p: ModelParameters, kfs: AbstractKFaceScheme,
                    sl: AbstractSlidingLaw) -> None:
    """src/elliptic_solver.jl:127-129."""
    Picard_loop(ps, s, g, p, kfs, sl)


def parabolic_solver(ls: CGIterativeSolver, s: State, g: Grid, p: ModelParameters, kfs: AbstractKFaceScheme,
                     sl: AbstractSlidingLaw, dt: float) -> None:
    """src/parabolic_solver.jl:16-38."""
    s._bind(p=p, kfs=kfs, sl=sl)
    _bind_solver(s, ls)
    cgi = C.c_int64(0)
    L.check(L.load().shk_parabolic_step(s.ctx, float(dt), C.byref(cgi)), s.ctx, "shk_parabolic_step")
    ls.ws.stats.niter = cgi.value


# -------------------------------------------------------------- Simulation

class EllipticHeadScheme:
    def __init__(self, ps: PicardSolver):
        self.ps = ps


class ParabolicHeadScheme:
    def __init__(self, ls: AbstractLinearSolver):
        self.ls = ls


class Simulation:
    """src/simulation.jl:105-184, with the observer choices of :162-181
    (which_observer "IO" or "Live"; which_file_writer "CSV" or "NetCDF" --
    "HDF5" and "JLD2" need an HDF5 library this installation lacks)."""

    def __init__(self, grid: Grid, state: State, tsteps: int, dt: float, p: ModelParameters, gap_scheme_choice: str,
                 tracked_obs, mi: AbstractMeltInput, sl: AbstractSlidingLaw, ps: PicardSolver | None = None,
                 ls: AbstractLinearSolver | None = None, which_observer: str | None = None,
                 which_file_writer: str | None = None, tracked_times=None, path=None,
                 k_face_choice: str = "arithmetic", verbose: bool = False):
        tracked_obs = list(tracked_obs)
        for name in tracked_obs:
            if name not in L.FIELD_ID:
                raise L.ShaktiHIPError(f'Unknown tracked observable: "{name}" is not a field of State')
        if p.e_v == 0:
            if ps is None:
                raise L.ShaktiHIPError("ps (a PicardSolver) must be provided when p.e_v == 0 (elliptic head scheme)")
            self.hs = EllipticHeadScheme(ps)
        else:
            if ls is None:
                raise L.ShaktiHIPError("ls (an AbstractLinearSolver) must be provided when p.e_v != 0 "
                                       "(parabolic head scheme)")
            self.hs = ParabolicHeadScheme(ls)
        if gap_scheme_choice not in ("explicit", "implicit"):
            raise L.ShaktiHIPError(f'Unknown gap_scheme_choice: "{gap_scheme_choice}" '
                                   '(expected "explicit" or "implicit")')
        if k_face_choice == "arithmetic":
            self.kfs = Arithmetic()
        elif k_face_choice == "harmonic":
            self.kfs = Harmonic()
        else:
            raise L.ShaktiHIPError(f'Unknown k_face_choice: "{k_face_choice}" (expected "arithmetic" or "harmonic")')
        self.grid, self.state, self.tsteps, self.dt, self.p = grid, state, int(tsteps), float(dt), p
        self.gap_scheme = 0 if gap_scheme_choice == "implicit" else 1
        self.mi, self.sl, self.verbose = mi, sl, verbose
        self.total_time = 0.0
        if isinstance(mi, SeasonalMoulinMeltInput) and mi.moulin is not None:
            a = mi.moulin
            L.check(L.load().shk_set_moulin_field(state.ctx, a.ctypes.data_as(C.POINTER(C.c_double)), a.size),
                    state.ctx, "shk_set_moulin_field")
        self.step_times: list[float] = []
        self.picard_log: list[tuple] = []
        self.observer = _make_observer(tracked_obs, which_observer, which_file_writer, tracked_times, path)


def _make_observer(tracked_obs, which_observer, which_file_writer, tracked_times, path):
    """The observer set-up of Simulation (src/simulation.jl:148-181), same errors."""
    if not tracked_obs:
        return O.NoObserver()
    if which_observer is None:
        raise L.ShaktiHIPError('which_observer should be specified as "IO" or "Live" when tracked_obs is not empty.')
    if which_observer == "IO":
        writers = {"NetCDF": O.NetCDFFileWriter, "HDF5": O.HDF5FileWriter, "JLD2": O.JLD2FileWriter,
                   "CSV": O.CSVFileWriter}
        if which_file_writer is None:
            raise L.ShaktiHIPError('which_file_writer should be specified as "NetCDF", "HDF5", "JLD2", or "CSV" '
                                   'when which_observer is "IO".')
        if which_file_writer not in writers:
            raise L.ShaktiHIPError(f'Unknown which_file_writer: "{which_file_writer}" '
                                   '(expected "NetCDF", "HDF5", "JLD2", or "CSV")')
        fr = writers[which_file_writer]()
        if tracked_times is None:
            raise L.ShaktiHIPError('tracked_times should be specified when which_observer is "IO".')
        if path is None:
            raise L.ShaktiHIPError('path should be specified when which_observer is "IO".')
        return O.IOObserver(tracked_obs, tracked_times, fr, path)
    if which_observer == "Live":
        if tracked_times is None:
            raise L.ShaktiHIPError('tracked_times should be specified when which_observer is "Live".')
        return O.LiveObserver(tracked_obs, tracked_times)
    raise L.ShaktiHIPError(f'Unknown which_observer: "{which_observer}" (expected "IO" or "Live")')


def picard_status(hs):
    """src/run.jl:73-74."""
    if isinstance(hs, EllipticHeadScheme):
        return hs.ps.converged, hs.ps.last_iter
    return None, None


def step(sim: Simulation) -> None:
    """src/run.jl:82-88: update_ieb!, step_h!, step_b!."""
    s = sim.state
    s._bind(p=sim.p, kfs=sim.kfs, sl=sim.sl, gap_scheme=sim.gap_scheme)
    mi = sim.mi
    mode = 0
    u = 0.0
    if isinstance(mi, SeasonalMoulinMeltInput):
        mode, u = 2, mi.value(sim.total_time)
    elif isinstance(mi, SeasonalMeltInput):
        mode, u = 1, mi.value(sim.total_time)
    conv, last, cgi = C.c_int32(), C.c_int32(), C.c_int64(0)
    if isinstance(sim.hs, EllipticHeadScheme):
        ps = sim.hs.ps
        _bind_solver(s, ps.ls)
        L.check(L.load().shk_step(s.ctx, sim.dt, mode, u, ps.iters, ps.tol, ps.check_every,
                                  -1.0 if ps.alpha is None else float(ps.alpha), C.byref(conv), C.byref(last),
                                  C.byref(cgi)), s.ctx, "shk_step")
        ps.converged, ps.last_iter, ps.cg_iters = bool(conv.value), last.value, cgi.value
    else:
        _bind_solver(s, sim.hs.ls)
        L.check(L.load().shk_step(s.ctx, sim.dt, mode, u, 1, 0.0, 1, -1.0, C.byref(conv), C.byref(last),
                                  C.byref(cgi)), s.ctx, "shk_step")
        sim.hs.ls.ws.stats.niter = cgi.value


def save_checkpoint(path, sim: Simulation, t: int, asynchronous: bool = False) -> None:
    """save_checkpoint (src/checkpoint.jl:26-38): every State field, t and
    total_time, written to path + ".tmp" and renamed over path.  With
    asynchronous, the call returns once the fields are copied on the device;
    the copy to the host and the file write go on behind later steps, and
    wait_checkpoint (or the next save) reports their outcome."""
    L.check(L.load().shk_checkpoint_save(sim.state.ctx, os.fsencode(str(path)), int(t), float(sim.total_time),
                                         1 if asynchronous else 0), sim.state.ctx, "shk_checkpoint_save")


def wait_checkpoint(sim: Simulation) -> None:
    """Waits for an asynchronous save_checkpoint and raises if it failed."""
    L.check(L.load().shk_checkpoint_wait(sim.state.ctx), sim.state.ctx, "shk_checkpoint_wait")


def load_checkpoint(sim: Simulation, path) -> int:
    """load_checkpoint! (src/checkpoint.jl:47-64): restores every State field
    and sim.total_time in place (all or nothing, each field's checksum checked
    first) and returns the saved timestep."""
    t, tt = C.c_int64(), C.c_double()
    L.check(L.load().shk_checkpoint_load(sim.state.ctx, os.fsencode(str(path)), C.byref(t), C.byref(tt)),
            sim.state.ctx, "shk_checkpoint_load")
    sim.total_time = tt.value
    return int(t.value)


def run(sim: Simulation, checkpoint_every: int | None = None, checkpoint_path=None, restart_path=None) -> None:
    """run! (src/run.jl:21-63): steps from 1 (or from the checkpoint at
    restart_path) to sim.tsteps, total_time += dt after each, the observer,
    a checkpoint every checkpoint_every steps and the per-step report line of
    :51.  Checkpoints are written asynchronously (the next steps run while
    the previous one goes to disk); run returns once the last one is on disk."""
    if (checkpoint_every is None) != (checkpoint_path is None):
        raise L.ShaktiHIPError("checkpoint_every and checkpoint_path must be given together")
    if restart_path is None:
        sim.total_time = 0.0
        O.prepare(sim.observer, sim.state)
        O.observe(sim.observer, sim.state, 0, sim.total_time)
        start_t = 0
    else:
        start_t = load_checkpoint(sim, restart_path)
        O.resume(sim.observer, sim.state, start_t)
    for t in range(start_t + 1, sim.tsteps + 1):
        t0 = time.perf_counter()
        step(sim)
        dt_wall = time.perf_counter() - t0
        sim.total_time += sim.dt
        O.observe(sim.observer, sim.state, t, sim.total_time)
        if checkpoint_every is not None and t % checkpoint_every == 0:
            save_checkpoint(checkpoint_path, sim, t, asynchronous=True)
        sim.step_times.append(dt_wall)
        conv, last = picard_status(sim.hs)
        sim.picard_log.append((conv, last))
        if sim.verbose:
            c = "missing" if conv is None else str(conv).lower()
            li = "missing" if last is None else last
            print(f"{t} / {sim.tsteps} completed in {round(dt_wall, 4)}s. Picard converged: {c} in {li} iterations",
                  flush=True)
    if checkpoint_every is not None:
        wait_checkpoint(sim)
    O.finalize(sim.observer, sim.state)


def abi_version() -> int:
    return L.load().shk_abi_version()
