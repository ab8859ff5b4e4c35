"""shakti_hip -- MI355X (gfx950) hydraulic-head solver for Shakti.jl's
linear-solver plugin seam, driven through the C ABI of libshakti_hip.so."""
from ._lib import LIB_PATH, ShaktiHIPError, load  # noqa: F401
from .api import (  # noqa: F401
    GROUNDED, LAND, OCEAN, OTHER_BASIN, Arithmetic, CGIterativeSolver, ConstantMeltInput, EllipticHeadScheme, Grid,
    Harmonic, HIPMatrixFreeLinearSystem, LinearSlidingLaw, MatrixFreeLinearSystem, ModelParameters,
    ParabolicHeadScheme, PicardSolver, Picard_loop, PrescribedSlidingLaw, RegularizedCoulombFieldSlidingLaw,
    RegularizedCoulombSlidingLaw, RegularizedCoulombV0SlidingLaw, SeasonalMeltInput,
    SeasonalMoulinMeltInput, Simulation, SparseAssembledLinearSystem, State, elliptic_solver, parabolic_solver,
    load_checkpoint, picard_status, run, save_checkpoint, set_initial_conditions, set_initial_conditions_on_device,
    solve_elliptic_linear_system,
    solve_parabolic_linear_system, step, update_chebyshev_bounds, update_derived, wait_checkpoint)
from .observer import (  # noqa: F401
    CSVFileWriter, HDF5FileWriter, IOObserver, JLD2FileWriter, LiveObserver, NetCDFFileWriter, NoObserver)
