/*
 * shakti_hip.h -- C ABI of libshakti_hip.so, the MI355X (gfx950) hydraulic-head
 * solver that stands behind Shakti.jl's linear-solver plugin seam.
 *
 * The library owns a device-resident copy of the whole State (every field of
 * the reference's src/state.jl:13-57, same column-major ix-fastest layout) and
 * the solver workspace.  Host code moves fields in and out with shk_upload /
 * shk_download and calls the entry points below; each one does the same work,
 * in the same order, as the reference function it names, on the GPU.
 *
 * Conventions
 *   - Every function returns SHK_OK (0) or a negative SHK_ERR_* code; the text
 *     of the last error of a context is returned by shk_last_error().  Nothing
 *     throws across the ABI.
 *   - One context is used by one host thread at a time; calls on a context are
 *     serialised on the context's own HIP stream and return after that stream
 *     has synchronised (except the *_async benchmarking helpers, which say so).
 *   - Host buffers belong to the caller and are copied.  All device memory
 *     belongs to the context and is released by shk_destroy().
 *   - Indexing: centre fields are nx*ny doubles (index ix + iy*nx), x-face
 *     fields (nx+1)*ny (index ix + iy*(nx+1)), y-face fields nx*(ny+1)
 *     (index ix + iy*nx).
 */
#ifndef SHAKTI_HIP_H
#define SHAKTI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHK_ABI_VERSION 2

#define SHK_OK 0
#define SHK_ERR_ARG (-1)      /* bad argument, shape or field id */
#define SHK_ERR_HIP (-2)      /* a HIP runtime call failed */
#define SHK_ERR_NUMERIC (-3)  /* breakdown: r'z < 0 or not finite (preconditioner or operator not positive
                                 definite; Krylov's sqrt(r'z) throws there), p'Ap or the returned head's
                                 residual not finite */
#define SHK_ERR_STATE (-4)    /* call out of order (e.g. solve before init) */
#define SHK_ERR_NODEVICE (-5) /* no usable gfx950 device */

/* Grid geometry -- reference src/grid.jl:8-35 (dx = lx/(nx-1), dy = ly/(ny-1)). */
typedef struct {
    int32_t nx, ny;
    double lx, ly, dx, dy, dx2, dy2;
} shk_grid;

/* Physical parameters -- reference src/model_parameters.jl:9-29 (defaults :55-84).
 * The Glen exponent n is canonicalised as the reference does
 * (src/model_parameters.jl:122-134): integer-valued n, n-1 and 1/n use the
 * integer power path, anything else the general power. */
typedef struct {
    double rho_w, rho_sw, rho_i, g, nu, n, omega, L, br, lr, ct, cw, p_atm, b_min, b_max, e_v;
} shk_params;

/* Sliding laws (src/sliding_law.jl).  The face stress of each: */
#define SHK_SLIDING_REGULARIZED_COULOMB 0 /* N C (v / (v + |N|^n lambda))^(1/n) ub / v, scalar C (:33, :213-249) */
#define SHK_SLIDING_PRESCRIBED 1          /* taub as uploaded before set-up, never recomputed (:45, :103-107) */
#define SHK_SLIDING_LINEAR 2              /* C^2 N ub, C a field (:61-90, :552-602) */
#define SHK_SLIDING_REGULARIZED_COULOMB_FIELD 3 /* the regularised-Coulomb stress with a field C (:270-396) */
#define SHK_SLIDING_REGULARIZED_COULOMB_V0 4    /* N C (v / (v + v0))^(1/n) ub / v, C a field (:422-543) */

/* Scheme choices of a Simulation -- reference src/simulation.jl:105-146,
 * src/sliding_law.jl.  Sensible heat is on iff ct*cw != 0 and opening by
 * sliding iff br != 0, decided from shk_params as the reference does.  The
 * field laws (LINEAR, _FIELD, _V0) take their cell-centred coefficient from
 * shk_set_sliding_coefficient. */
typedef struct {
    int32_t kfs;         /* 0 Arithmetic, 1 Harmonic (src/k_face_scheme.jl:15,23) */
    int32_t gap_scheme;  /* 0 implicit, 1 explicit (src/run.jl:130-145) */
    int32_t sliding_law; /* SHK_SLIDING_* */
    double sliding_C;    /* RegularizedCoulombSlidingLaw.C */
    double sliding_v0;   /* RegularizedCoulombV0SlidingLaw.v0 (m/s) */
} shk_model;

#define SHK_PRECOND_JACOBI 0    /* CGIterativeSolver(g, MatrixFreeLinearSystem) */
#define SHK_PRECOND_CHEBYSHEV 1 /* ... ; chebyshev_degree = k */
#define SHK_PRECOND_MULTIGRID 2 /* geometric multigrid V-cycle in FP32: with SHK_FORM_FP64 it preconditions
                                   FP64 CG, with SHK_FORM_FP32_REFINE the FP32 inner solves; no reference
                                   counterpart (the reference's AMG is CPU-only) */

#define SHK_FORM_FP64 0         /* plain FP64 PCG */
#define SHK_FORM_FP32_REFINE 1  /* FP32 inner PCG inside FP64 iterative refinement */
#define SHK_FORM_LEARNED 2      /* FP64 PCG warm-started by the attached U-Net (shakti_learned.h) */

/* Chebyshev bounds.  REFERENCE: 0.9 x the smallest and 1.1 x the largest Ritz
 * value of the nsteps-step estimate with Euclidean dots on D^-1 A, adopted
 * when finite with 0 < min < max, else the previous bounds
 * (src/preconditioner.jl:89-134, 197-233).  Those Ritz values are not
 * bounded by the spectrum of the nonsymmetric D^-1 A: on strongly varying
 * transmissivity the estimate can be off by many orders of magnitude, and
 * the shrunken polynomial then trips Krylov's zero-curvature test.
 * GERSHGORIN (default): the same estimate run in the D-inner product (the
 * coefficients of Jacobi-preconditioned CG, a symmetric Lanczos process, so
 * its Ritz values lie in the spectrum) gives lambda_min = 0.9 x the smallest;
 * lambda_max = the Gershgorin bound of D^-1 A, 1 + max_i sum_j |a_ij| / a_ii
 * (at most 2 for these diagonally dominant systems).  The reference's
 * degree-k polynomial is positive exactly below lambda_min + lambda_max, so
 * with this lambda_max the preconditioner is positive definite.  A deviation
 * from the reference, which it follows otherwise. */
#define SHK_CHEB_BOUNDS_GERSHGORIN 0
#define SHK_CHEB_BOUNDS_REFERENCE 1

/* shk_solver_cfg.flags.  PICARD_ATOL: inside a Picard loop, every solve after
 * the first stops at atol + rtol * rNorm0 with atol raised to rtol times the
 * first solve's initial residual norm, so a later solve, which starts from a
 * head already close to the fixed point, is held to the first solve's
 * absolute target instead of reducing its own small residual by rtol again.
 * No reference counterpart (Krylov's atol is a fixed number); off (0) by
 * default. */
#define SHK_CFG_PICARD_ATOL 1

/* Linear-solver configuration -- reference CGIterativeSolver(g,
 * MatrixFreeLinearSystem; chebyshev_degree, chebyshev_nsteps_estimate)
 * (src/linear_solver.jl:832-843) plus Krylov's stopping controls. */
typedef struct {
    int32_t precond;          /* SHK_PRECOND_* */
    int32_t cheb_degree;      /* >= 1 */
    int32_t cheb_nsteps;      /* Lanczos steps for the bound estimate, >= 2 (default 15) */
    int32_t form;             /* SHK_FORM_* */
    int32_t itmax;            /* 0 -> 2*nx*ny, Krylov's default */
    int32_t check_every;      /* CG iterations queued between host convergence checks (>= 1) */
    double atol, rtol;        /* stop when sqrt(r'M^-1 r) <= atol + rtol*initial; Krylov default sqrt(eps) each */
    double inner_rtol;        /* FP32_REFINE: relative reduction asked of each inner solve */
    int32_t cheb_bounds;      /* SHK_CHEB_BOUNDS_* (ABI 2) */
    int32_t flags;            /* SHK_CFG_* bits (ABI 2) */
} shk_solver_cfg;

/* Statistics of the last linear solve (reference ws.stats, ls.precond bounds). */
typedef struct {
    int32_t niter;          /* CG iterations (all inner iterations for FP32_REFINE) */
    int32_t solved;         /* 1 when the stopping test was met */
    int32_t zero_curvature; /* 1 when p'Ap vanished (Krylov's zero-curvature exit) */
    int32_t bounds_warning; /* 1 when the Chebyshev bound estimate was rejected (previous bounds kept) */
    int32_t refinements;    /* FP32_REFINE outer iterations */
    int32_t fallback;       /* LEARNED: bit 0 when the network's warm start was rejected (CG started
                               from h); bit 1 when the learned preconditioner stalled (plain PCG after) */
    double rnorm0, rnorm;   /* preconditioned residual norm at start / end */
    double lambda_min, lambda_max; /* Chebyshev bounds in use */
    /* ABI 2 */
    int32_t tired;          /* 1 when itmax was reached before the stopping test (FP32_REFINE: 64 refinements;
                               a refinement loop that stalls at the FP64 floor, three refinements without
                               a 10% gain, ends with solved = 0 and tired = 0) */
    int32_t breakdown;      /* 0; 1 when r'z < 0 or not finite, 2 when p'Ap is not finite (the solve then
                               returned SHK_ERR_NUMERIC and left h unchanged) */
    double breakdown_value; /* the offending r'z or p'Ap */
    double lz_lambda_min, lz_lambda_max; /* Chebyshev: this solve's estimate, 0.9 x / 1.1 x the extreme Ritz
                                            values, whether adopted or not (NaN otherwise) */
    double gershgorin;      /* Chebyshev with SHK_CHEB_BOUNDS_GERSHGORIN: the bound of D^-1 A (NaN otherwise) */
    double true_residual;   /* ||b - A h||_2 / ||b||_2 of the head the solve returned, in FP64 (before any
                               under-relaxation) */
} shk_solve_stats;

/* State field ids, in the order of src/state.jl:13-57. */
enum shk_field {
    SHK_H = 0, SHK_PW, SHK_PO, SHK_B, SHK_BETA, SHK_ABS_UB, SHK_MDOT, SHK_SHEAR, SHK_POTENTIAL, SHK_SENSIBLE,
    SHK_RE, SHK_K, SHK_G, SHK_ZB, SHK_ZS, SHK_HICE, SHK_IEB, SHK_LAMBDA, SHK_A_VISC, SHK_N, SHK_MASK,
    SHK_DHDX, SHK_Q_X, SHK_RE_X, SHK_B_X, SHK_UB_X, SHK_TAUB_X, SHK_DPWDX, SHK_VALID_X,
    SHK_DHDY, SHK_Q_Y, SHK_RE_Y, SHK_B_Y, SHK_UB_Y, SHK_TAUB_Y, SHK_DPWDY, SHK_VALID_Y,
    SHK_NUM_FIELDS
};

typedef struct shk_ctx shk_ctx;

/* Library identification: ABI version and the device the context runs on.
 * shk_last_error(NULL) returns the message of the last failed context-free
 * call of the calling thread. */
int32_t shk_abi_version(void);
const char* shk_last_error(const shk_ctx* ctx);

/* Create / destroy a context on the current HIP device (allocates every State
 * field, zeroed, with mask = GROUNDED and valid_x/valid_y = 1 as State(g) does,
 * src/state.jl:86-137, plus the solver workspace).
 * Replaces: State(g) + CGIterativeSolver(g, MatrixFreeLinearSystem; ...)
 * (src/state.jl:86, src/linear_solver.jl:832). */
int shk_create(const shk_grid* grid, const shk_params* params, const shk_model* model, const shk_solver_cfg* cfg,
               shk_ctx** out);
int shk_destroy(shk_ctx* ctx);

/* Replace the solver configuration (keeps the persistent Chebyshev bounds). */
int shk_set_solver_cfg(shk_ctx* ctx, const shk_solver_cfg* cfg);

/* Replace the physical parameters / scheme choices used by later calls (the
 * reference passes p and kfs to every solve, src/linear_solver.jl:915). */
int shk_set_params(shk_ctx* ctx, const shk_params* params);
int shk_set_model(shk_ctx* ctx, const shk_model* model);

/* The cell-centred coefficient C (nx*ny values) of the field sliding laws,
 * staggered onto the faces on the device as their constructors do
 * (LinearSlidingLaw(g, C) src/sliding_law.jl:73-90,
 * RegularizedCoulombFieldSlidingLaw(g, C) :284-301,
 * RegularizedCoulombV0SlidingLaw(g, C, v0) :436-452): an outer face takes its
 * cell's value, an inner face the mean of its two cells.  The linear law
 * squares the face value where it uses it (the same product as the
 * reference's stored Cx^2). */
int shk_set_sliding_coefficient(shk_ctx* ctx, const double* C_host, size_t count);

/* Copy one State field to / from the device; count must equal its size. */
int shk_upload(shk_ctx* ctx, int32_t field, const double* host, size_t count);
int shk_download(shk_ctx* ctx, int32_t field, double* host, size_t count);

/* Checkpoint / restart: save_checkpoint(path, sim, t) and
 * load_checkpoint!(sim, path) (src/checkpoint.jl:26-64), which run! calls
 * every checkpoint_every steps and on restart (src/run.jl:21-47).
 *
 * shk_checkpoint_save writes every State field plus t and total_time to
 * path + ".tmp", flushes it to disk and renames it over path, so path always
 * holds a complete checkpoint.  The fields are first copied on the device
 * (the State may change as soon as the call returns); with async != 0 the
 * copy to the host and the file write then proceed on a second stream and a
 * host thread while later steps run, and shk_checkpoint_wait returns the
 * write's outcome (it is also awaited by the next save, a load and
 * shk_destroy).  The file is the library's own format: a 64-byte header
 * (magic "SHKCKPT1", version 1, field count, nx, ny, t, total_time, lx, ly,
 * data offset), one 64-byte record per field (the reference State's field
 * name, id, rows, cols, offset, bytes, checksum), then each field in FP64,
 * column-major, 512-byte aligned.  The checksum of n values with bit patterns
 * w_i is sum_i splitmix64(w_i + i * 0x9E3779B97F4A7C15) mod 2^64.
 *
 * shk_checkpoint_load checks the header, the grid, every field's shape and
 * checksum before it replaces any field (all or nothing), then returns t and
 * total_time.  A failed check leaves the State as it was and returns
 * SHK_ERR_ARG (unreadable, wrong grid or shape) or SHK_ERR_NUMERIC (checksum). */
int shk_checkpoint_save(shk_ctx* ctx, const char* path, int64_t t, double total_time, int32_t async);
int shk_checkpoint_wait(shk_ctx* ctx);
int shk_checkpoint_load(shk_ctx* ctx, const char* path, int64_t* t, double* total_time);

/* Asynchronous field output for observers (src/observer.jl:263-348 reads the
 * tracked fields at each tracked time).  shk_snapshot copies the listed
 * fields (all of them for field_ids NULL or nfields 0) on the device and
 * queues their copy to pinned host memory on a second stream, so the caller
 * can go on stepping; shk_snapshot_fetch waits for that copy and hands out one
 * field of the last snapshot. */
int shk_snapshot(shk_ctx* ctx, const int32_t* field_ids, int32_t nfields);
int shk_snapshot_fetch(shk_ctx* ctx, int32_t field, double* host, size_t count);

/* min, max and mean of one State field, reduced on the device: the row the
 * CSV observer writes per tracked field (extrema and mean,
 * src/observer.jl:324-339); NaN anywhere gives NaN for all three. */
int shk_field_stats(shk_ctx* ctx, int32_t field, double* out3);

/* set_initial_conditions! (src/initial_conditions.jl:14-126) on the device.
 * The raw inputs mask, A_visc, zb, zs, b, G, ub_x, ub_y, ieb (and taub_x,
 * taub_y for the prescribed law) must have been uploaded first. */
int shk_set_initial_conditions(shk_ctx* ctx);

/* The synthetic glacier of the benchmark configurations, generated on the
 * device: writes the raw inputs of shk_set_initial_conditions (mask, A_visc,
 * zb, zs, b, G, ub_x, ub_y, ieb, taub_x, taub_y) and, with n_moulins > 0, the
 * moulin field of ieb mode 2.  No reference counterpart: the reference's
 * examples build these inputs on the host (src/initial_conditions.jl:14-126
 * takes them as arguments); this removes the host construction and upload of
 * large grids.  Definition, reproducible bit for bit by a CPU restatement:
 *   stream  o_i = SplitMix64 output i from state seed;
 *   noise   w_c = sum_{k<12} (o_{12c+k} >> 16) 2^-48 - 6, c = ix + iy nx;
 *   bed     periodic Gaussian smoothing along x then y, weights
 *           exp(-(k/s)^2/2) normalised over k = -R..R, s = corr_len /
 *           (sqrt(2) dx) (dy along y), R = ceil(4 s) (none for corr_len 0),
 *           so the covariance is exp(-r^2 / (2 corr_len^2)); then shifted to
 *           zero mean and scaled to standard deviation bed_amp; minus a
 *           Gaussian trough depth exp(-((y - y0) / (width/2))^2 / 2),
 *           y0 = (1/4 + u/2) ly, when trough_width > 0;
 *   zs      6 (sqrt(x + 5000) - sqrt(5000)) + 1, x = ix dx;
 *   mask    GROUNDED except the ix = 0 column (OCEAN); A_visc, b = b0, G,
 *           ieb uniform; ub_x = ub, ub_y = 0, taub = 0;
 *   moulins n_moulins distinct interior cells, Q ~ U[q_min, q_max] m^3/s,
 *           moulin = Q / (dx dy) there and 0 elsewhere.
 * Requires nx, ny >= 2, and nx, ny >= 3 for moulins. */
typedef struct {
    uint64_t seed;
    double bed_amp, corr_len, trough_width, trough_depth;
    double b0, A_visc, ub, G, ieb;
    int32_t n_moulins, reserved;
    double q_min, q_max;
} shk_glacier_spec;
int shk_generate_glacier(shk_ctx* ctx, const shk_glacier_spec* spec);

/* solve_elliptic_linear_system!(ls, s, g, p, kfs) for the matrix-free CG
 * solver (src/linear_solver.jl:915-925): assemble, refresh Jacobi and the
 * Chebyshev bounds, PCG warm-started from h; overwrites device h. */
int shk_solve_elliptic(shk_ctx* ctx, shk_solve_stats* stats);

/* solve_parabolic_linear_system!(ls, s, g, p, kfs, dt) (src/linear_solver.jl:935-945). */
int shk_solve_parabolic(shk_ctx* ctx, double dt, shk_solve_stats* stats);

/* The field refresh that follows every linear solve: dhdxy, pw, dpwdxy, N,
 * q_and_Re_xy, Re, taub_xy, mdot, K (src/elliptic_solver.jl:193-206). */
int shk_update_derived(shk_ctx* ctx);

/* One Picard_iteration! (src/elliptic_solver.jl:187-208); relax_alpha <= 0
 * means NoHeadRelaxation. */
int shk_picard_iteration(shk_ctx* ctx, double relax_alpha, shk_solve_stats* stats);

/* Picard_loop! (src/elliptic_solver.jl:138-176): converged / last_iter as the
 * reference's ps.converged / ps.last_iter; cg_iters accumulates CG iterations.
 * An iteration whose linear solve stopped short (itmax or zero curvature:
 * stats tired / zero_curvature) does not count as converged, whatever the
 * head change (a deviation: the reference tests the head change alone).  A
 * breakdown returns SHK_ERR_NUMERIC. */
int shk_picard_loop(shk_ctx* ctx, int32_t iters, double tol, int32_t check_every, double relax_alpha,
                    int32_t* converged, int32_t* last_iter, int64_t* cg_iters);

/* Statistics of every linear solve since the last call (oldest first), for
 * the solves inside shk_picard_loop / shk_step as well.  Copies at most
 * max_entries (the most recent ones when there are more), sets *n_entries and
 * empties the log, which keeps the last 65536 solves.  No reference
 * counterpart beyond ls.ws.stats of the last solve. */
int shk_solve_log(shk_ctx* ctx, shk_solve_stats* out, int32_t max_entries, int32_t* n_entries);

/* parabolic_solver! (src/parabolic_solver.jl:16-38). */
int shk_parabolic_step(shk_ctx* ctx, double dt, int64_t* cg_iters);

/* step_b! (src/run.jl:154-164): compute_b! by the context's gap scheme, then
 * beta (when br != 0), b_x, b_y. */
int shk_step_b(shk_ctx* ctx, double dt);

/* update_ieb! (src/melt_input.jl:30, 86-94): mode 0 leaves ieb untouched
 * (ConstantMeltInput), mode 1 overwrites it with `uniform` (SeasonalMeltInput),
 * mode 2 writes uniform + the static moulin field set by shk_set_moulin_field
 * (point moulins on top of the seasonal input; no reference counterpart). */
int shk_set_moulin_field(shk_ctx* ctx, const double* host, size_t count);
int shk_update_ieb(shk_ctx* ctx, int32_t mode, double uniform);

/* One whole step! (src/run.jl:82-88): update_ieb!, step_h! (Picard when
 * e_v == 0, else the parabolic solve), step_b!. */
int shk_step(shk_ctx* ctx, double dt, int32_t ieb_mode, double ieb_uniform, int32_t iters, double tol,
             int32_t check_every, double relax_alpha, int32_t* converged, int32_t* last_iter, int64_t* cg_iters);

/* Stencil operator apply y = A x with the coefficients of the last assembly
 * (StencilOperator mul!, src/linear_solver.jl:616-669).  Host buffers of
 * nx*ny doubles. */
int shk_assemble(shk_ctx* ctx, int32_t parabolic, double dt);
int shk_apply_operator(shk_ctx* ctx, const double* x_host, double* y_host);
/* update_chebyshev_bounds!(P, rhs) (src/preconditioner.jl:197-233) with the
 * estimate of src/preconditioner.jl:89-134 run on the operator of the last
 * assembly from a host rhs of nx*ny doubles, under the configuration's
 * bound mode and step count (the configured preconditioner must be
 * Chebyshev).  A valid estimate replaces the persistent bounds; an invalid
 * one (e.g. from an all-zero rhs) keeps them, the first time the placeholder
 * (1, 1), and sets *warned.  Any output pointer may be NULL. */
int shk_update_chebyshev_bounds(shk_ctx* ctx, const double* rhs_host, double* lambda_min, double* lambda_max,
                                int32_t* warned);
/* ||b - A x||_2 and ||b||_2 of the last assembly in FP64 (norms[0], norms[1])
 * for a host vector x of nx*ny doubles: the true residual of a head, formed
 * with the operator's own apply and two sums of squares finished in a fixed
 * order.  Krylov's cg! and the reference keep only the recursive residual;
 * this is the check the solves report as stats.true_residual. */
int shk_residual_norms(shk_ctx* ctx, const double* x_host, double* norms);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The result of shk_residual_norms, bit for bit, for an
 * FP64 system given in device memory in the symmetric face layout (aP, aE,
 * aN, b, x: nx*ny doubles each, ix fastest; aE couples (ix, iy) to
 * (ix+1, iy), aN to (ix, iy+1)): ||b - A x||_2 into out_dev[0] and ||b||_2
 * into out_dev[1] (device memory).  Enqueued on `stream` (a hipStream_t;
 * NULL is the null stream) and asynchronous: no allocation, no copy, no
 * synchronisation.  The per-workgroup partials go to one static device
 * array of the library, 512 KiB per GPU, held until the process ends; grids
 * that need more than 32768 workgroups (above about 1.3e8 cells for even nx,
 * 2.7e8 for odd nx) are
 * rejected with SHK_ERR_ARG.  Every call on a GPU shares that array, so
 * calls that overlap on one GPU race.  Takes no context: on failure
 * shk_last_error(NULL) gives the message (per host thread). */
int shk_residual_norms_noalloc(int32_t nx, int32_t ny, const double* aP, const double* aE, const double* aN,
                               const double* b, const double* x, double* out_dev, void* stream);
/* One application z = M r of the multigrid preconditioner (the configured
 * one must be SHK_PRECOND_MULTIGRID) built from the last assembly, as the
 * solve applies it: one V-cycle of the FP32 flux-form hierarchy with an FP64
 * r and z.  Host buffers of nx*ny doubles.  No counterpart in the reference
 * (whose preconditioners are Jacobi, Chebyshev and CPU-side AMG); for
 * callers that run their own Krylov loop, and for checking the V-cycle. */
int shk_apply_precond(shk_ctx* ctx, const double* r_host, double* z_host);
/* Download the last assembly in the reference's MatrixFreeLinearSystem layout
 * (aP, aE, aW, aN, aS, rhs; each nx*ny). Any pointer may be NULL. */
int shk_download_system(shk_ctx* ctx, double* aP, double* aE, double* aW, double* aN, double* aS, double* rhs);

/* Benchmark helpers: time `reps` back-to-back launches of one kernel on the
 * context's stream with HIP events; returns the mean duration in ms and the
 * algorithmic bytes one launch moves.  kernel: 0 FP64 operator apply,
 * 1 FP32 operator apply, 2 plain FP64 copy of nx*ny values (the streaming
 * reference point); 3 and 4 the same as 0 and 2 but cold: each launch is
 * timed alone after a 1 GiB read has replaced the contents of L2 and the
 * Infinity Cache (allocates that buffer for the duration of the call); 5 the
 * FP64 CG matvec kernel (p = z + beta p_old, Ap, p'Ap partials, and the
 * previous iteration's x += alpha p_old, here with alpha = 0 so x keeps its
 * values) on the solve's work vectors, in the variant the
 * V-cycle-preconditioned solve runs (z read in FP32 where that solve keeps it
 * so: 68 instead of 72 bytes per cell; where it also takes the x steps in
 * pairs, the odd iteration's kernel, without x, and the even one's, with two
 * steps, in turn: 52 and 76, a mean of 64), cold like 3 and 4: in that solve the
 * V-cycle between two matvecs streams their inputs out of the caches (the
 * device scalars are restored afterwards). */
int shk_time_kernel(shk_ctx* ctx, int32_t kernel, int32_t reps, double* mean_ms, double* bytes_per_launch);

/* Per-kernel timing collected with HIP events on the context's stream while
 * enabled: names[i] (static strings), launch counts, total milliseconds and
 * algorithmic bytes.  mask: 0 off, 1 every kernel, otherwise a bit set of
 * kernel ids (bit i = entry i of shk_profile_read) so that only the kernels of
 * interest carry event records. */
int shk_profile_enable(shk_ctx* ctx, int32_t mask);
int shk_profile_read(shk_ctx* ctx, int32_t max_entries, const char** names, int64_t* counts, double* total_ms,
                     double* bytes, int32_t* n_entries);
int shk_profile_reset(shk_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SHAKTI_HIP_H */
