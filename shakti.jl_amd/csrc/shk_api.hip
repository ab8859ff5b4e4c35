// C-ABI implementation of libshakti_hip.so (declared in include/shakti_hip.h).
//
// A context owns one HIP stream (default priority), every State field on the
// device, the linear-system arrays and the solver workspace.  The solver loops
// run on the host but never wait on the device per CG iteration: every kernel
// reads the device-resident scalars (alpha, beta, done, ...) written by the
// finishing workgroup of the previous reduction, kernels queued after
// convergence return at once, and the host looks at the `done` flag only every
// `check_every` iterations.

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "shakti_learned.h"
#include "shk_internal.hpp"
#include "shk_io.hpp"

using namespace shk;

// U-Net entry points used by the solver (shk_nn.hip)
namespace shknn {
int forward(shk_unet* net, const void* x, int B, int H, int W);
float* output(shk_unet* net);
void set_stream(shk_unet* net, hipStream_t st);
void release_stream(shk_unet* net, hipStream_t st);
bool fits(const shk_unet* net, int H, int W);
const char* unet_error(const shk_unet* net);
}  // namespace shknn

namespace {

struct ProfEntry {
    const char* name;
    int64_t count = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct Pending {
    int id;
    hipEvent_t a, b;
    double bytes;
};

}  // namespace

struct shk_ctx {
    shk_grid g{};
    shk_params p{};
    shk_model m{};
    shk_solver_cfg cfg{};
    Phys P{};
    Geom G{};
    hipStream_t st = nullptr;
    // the coarsest multigrid inverse is formed on st_side, beside the work
    // that precedes the first V-cycle tail (one workgroup would hold the
    // whole chip on st); st waits on inv_done before the next tail
    hipStream_t st_side = nullptr;
    hipEvent_t inv_fork = nullptr, inv_done = nullptr;
    bool inv_pending = false;
    int device = 0;
    size_t n = 0, nxf = 0, nyf = 0;

    Fields F{};
    double* h_alt = nullptr;  // ping-pong partner of the h field
    double* moulin = nullptr;
    // face coefficient of the field sliding laws (shk_set_sliding_coefficient)
    double *slide_cx = nullptr, *slide_cy = nullptr;
    // snapshots, checkpoints and field statistics (shk_io.hip), created on first use
    shkio::Snapshot* io = nullptr;
    // a lower bound on atol for the solves of a Picard loop (SHK_CFG_PICARD_ATOL)
    double atol_floor = 0.0;

    // linear system (FP64) and workspace
    double *aP = nullptr, *aE = nullptr, *aN = nullptr, *rhs = nullptr;
    double *x = nullptr, *r = nullptr, *z = nullptr, *pA = nullptr, *pB = nullptr, *Ap = nullptr;
    double *s = nullptr, *cA = nullptr, *cB = nullptr, *r2 = nullptr, *cP = nullptr;
    // FP32 copies for the mixed-precision form
    float *aP32 = nullptr, *aE32 = nullptr, *aN32 = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *z32 = nullptr, *pA32 = nullptr, *pB32 = nullptr,
          *Ap32 = nullptr, *s32 = nullptr, *cA32 = nullptr, *cB32 = nullptr, *r2_32 = nullptr, *dg32 = nullptr;
    double* d64 = nullptr;  // FP64 correction / residual scratch

    double* partials = nullptr;
    double* partials2 = nullptr;  // p'Ap partials while r'z partials are pending (deferred finishing)
    Scalars* sc = nullptr;       // device
    Scalars* sc_host = nullptr;  // pinned host mirror
    bool have_system = false;
    int system_parabolic = 0;
    double system_dt = 0.0;

    // profiling (HIP events around the launches whose id bit is set in prof_mask)
    bool prof = false;
    uint32_t prof_mask = 0xffffffffu;
    std::vector<ProfEntry> prof_tab;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;

    // geometric multigrid levels (FP32), allocated on first use
    struct MGLevel {
        int nx, ny;
        float *c, *e, *n, *D, *b, *xa, *xb;
        bool owned;
        bool packed = false;  // c, e, n, D live in mg_pack (tail levels)
        bool lin = false;     // reaches the next level by the linear transfers
    };
    std::vector<MGLevel> mg;
    float* mg_pack = nullptr;  // the tail's [c | e | n | D] (stride mg_stride) and the coarsest inverse
    int mg_stride = 0;
    float* mg_inv = nullptr;   // inside mg_pack
    // Smoother, its weight and the coarse-correction scale.  Red-black
    // Gauss-Seidel (the default) with over-relaxation 1.2: on the assembled
    // glacier systems 35-40% fewer CG iterations than damped Jacobi (omega
    // 0.95, scale 1.5; SHK_MG_JACOBI=1) at the same memory traffic per
    // V-cycle.  The piecewise-constant coarse correction is scaled up to make
    // up for the energy it misses, which depends on how diagonally dominant
    // the system is: 1.9 for the elliptic systems (the reaction term orders
    // of magnitude below the conductances; with two sweeps on every level,
    // 4096 x 1024 CG iterations per step at 1.6 / 1.8 / 1.9 / 1.95 / 2.0:
    // 124 / 116 / 112 / 111 / 110, the ragged test systems within one
    // iteration), 1.0 (no over-correction) for the backward-Euler parabolic
    // ones, where the storage term dominates (on the 8192 x 2048 workload,
    // CG iterations per solve for scale 0.8 / 0.9 / 1.0 / 1.1 / 1.2 / 1.8:
    // 50 / 42 / 37 / 41 / 49 / 150).  The symmetric V-cycle (red-black down,
    // black-red up) stays positive definite for omega < 2 and scale < 2
    // (with an exact coarse solve the correction I - s pi, pi a projection,
    // is a contraction exactly for 0 < s < 2).
    int mg_rb = 1;
    float mg_omega = 1.2f, mg_scale = 1.9f, mg_scale_par = 1.0f;
    // Transfers from level mg_lin_from down (red-black smoothing only): the
    // cell-centred linear ones, R = P^T, with the lumped Galerkin coarse
    // operator, in place of the 2 x 2 aggregation; their coarse correction
    // needs no over-correction (scale 0.8 elliptic, 1.0 parabolic).  On the
    // assembled glacier systems (1024 x 256, FP32/BF16 V-cycle, CG to 1e-10)
    // 22 CG iterations against aggregation's 34 with the finest transfer
    // still aggregating, 21 with every transfer linear.
    int mg_lin_from = 1;
    // smoothing sweeps each way on the levels below the finest that take the
    // two-row legs with the linear transfers (1 or 2; SHK_MG_NU), and on the
    // finest level where it takes the BF16 legs (SHK_MG_NU0)
    int mg_nu = 2, mg_nu0 = 2;
    float mg_scale_lin = 0.8f, mg_scale_lin_par = 1.0f;
    // The finest level's c, aE, aN, D in BF16 for its red-black legs (where
    // those are the two-column kernels): half the coefficient bytes of the
    // two full-size V-cycle kernels, at no cost in CG iterations measured
    // (SHK_MG_FP32_FINE=1 keeps FP32).  The FP32 arrays stay for coarsening.
    bool mg_bf16 = false;
    uint16_t *mg_c16 = nullptr, *mg_e16 = nullptr, *mg_n16 = nullptr, *mg_D16 = nullptr;
    size_t mg_tail_first = 1;                 // first level handled by the single-workgroup tail kernel
    size_t mg_tail_h = 0;                     // the tail's linear-restriction staging (floats)

    // captured CG iteration batches, keyed by (element size, jacobi, degree, batch)
    struct GraphEntry {
        int key[4];
        hipGraph_t graph;
        hipGraphExec_t exec;
        const shk_unet* net = nullptr;  // set for a batch that runs the attached network
    };
    std::vector<GraphEntry> graphs;
    bool use_graphs = true;
    bool use_fused = true;
    bool use_defer = true;  // FP64 CG + V-cycle: finishing kernels folded into the next kernel
    bool use_adaptive = true;  // batch sizes from the observed convergence rate
    int last_niter = 0;        // CG iterations of the previous solve
    int batch_margin = 0;      // the first batch stops this many iterations short of the prediction
    // CG iterations of the FP64 solve of each Picard iteration of the last loop
    // (the first batch of the same iteration of the next loop is sized from it)
    std::vector<int> niter_by_picard;
    int picard_it = 0;  // 1-based Picard iteration of the solve under way, 0 outside a loop
    // the Picard loop asks the solve to take the Picard test in its epilogue
    // (picard_fuse); the solve says whether it did (picard_fused)
    bool picard_fuse = false, picard_fused = false;
    // derived-field refresh: one marching pass (or the two-pass kernels); a
    // Picard iteration that another one follows refreshes only N, K and mdot
    bool derived_two_pass = false;
    bool lean_refresh = true;
    // the FP64 multigrid solve from h has its CG start (r = b - A h, x = 0,
    // p = 0) formed by the assembly
    bool asm_start = true;
    // FP64 CG with the V-cycle's FP32 z: the matvecs take the x steps in
    // pairs (x read and written every other iteration)
    bool pair_x = true;

    // learned warm start and training-pair recording: the network input is
    // the grid padded with zero cells to W8 x H8 (multiples of 64 and 32)
    shk_unet* net = nullptr;
    int W8 = 0, H8 = 0;
    void* nn_x = nullptr;  // BF16 network input [H8][W8][8]
    double *nn_e = nullptr, *nn_Ae = nullptr, *hw = nullptr;
    void* rec_x = nullptr;
    float* rec_t = nullptr;
    int rec_cap = 0, rec_count = 0;
    // rec_kind 0: one pair per Picard iteration; 1: one pair per learned CG
    // iteration, whose target needs the iterate there (rec_snap, n per slot)
    // and the image scale (rec_scale)
    int rec_kind = 0;
    double *rec_snap = nullptr, *rec_scale = nullptr;
    // learned preconditioner inside flexible CG: one iteration in nn_every
    // (0: warm start only); a learned batch whose residual norm does not fall
    // below nn_stall x its value at the batch start ends the learned batches
    int nn_every = 0;
    double nn_stall = 1.0;

    // statistics of the solves since the last shk_solve_log (at most kLogCap)
    std::vector<shk_solve_stats> solve_log;

    std::string err;
};

namespace {

const char* kProfNames[] = {"assemble",   "lanczos",      "cheb_bounds", "cg_init",     "cheb_pass",
                            "cg_matvec",  "cg_update",    "cg_finish",   "derived",     "gap",
                            "stagger_b",  "ieb",          "ic",          "matvec",      "ir",
                            "cg_update_cheb", "mg_down",  "mg_up",       "mg_tail",     "mg_setup",
                            "nn_forward", "nn_warm_start", "mg_down_fine", "mg_up_fine", "nn_precond"};
// mg_down / mg_up: the coarser levels; *_fine: the finest level
enum ProfId { P_ASM, P_LZ, P_CB, P_INIT, P_CHEB, P_MV, P_UPD, P_FIN, P_DER, P_GAP, P_STAG, P_IEB, P_IC, P_MATVEC, P_IR,
              P_FUSED, P_MGD, P_MGU, P_MGC, P_MGS, P_NNF, P_NNW, P_MGD0, P_MGU0, P_NNP, P_NUM };

constexpr size_t kLogCap = 65536;

int fail(shk_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return fail(ctx, SHK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipEvent_t take_event(shk_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Profiling brackets around one launch (no-ops unless enabled).
struct Prof {
    shk_ctx* c;
    int id;
    double bytes;
    hipEvent_t a = nullptr;
    Prof(shk_ctx* c_, int id_, double bytes_) : c(c_), id(id_), bytes(bytes_) {
        if (c->prof && ((c->prof_mask >> id) & 1u)) {
            a = take_event(c);
            if (a) hipEventRecord(a, c->st);
        }
    }
    ~Prof() {
        if (c->prof && a) {
            hipEvent_t b = take_event(c);
            if (!b) return;
            hipEventRecord(b, c->st);
            c->pending.push_back({id, a, b, bytes});
        }
    }
};

void prof_collect(shk_ctx* c) {
    if (c->pending.empty()) return;
    hipStreamSynchronize(c->st);
    for (auto& pe : c->pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, pe.a, pe.b);
        c->prof_tab[pe.id].count++;
        c->prof_tab[pe.id].ms += ms;
        c->prof_tab[pe.id].bytes += pe.bytes;
        c->ev_pool.push_back(pe.a);
        c->ev_pool.push_back(pe.b);
    }
    c->pending.clear();
}

Expo canon(double e) {
    Expo x;
    x.e = e;
    x.is_int = (e == (double)(long)e) && e > -64 && e < 64;
    x.k = x.is_int ? (int)e : 0;
    x.third = e == 1.0 / 3.0;
    return x;
}

void fill_phys(shk_ctx* c) {
    Phys& P = c->P;
    P.nx = c->g.nx; P.ny = c->g.ny;
    P.dx = c->g.dx; P.dy = c->g.dy; P.dx2 = c->g.dx2; P.dy2 = c->g.dy2;
    const shk_params& p = c->p;
    P.rho_w = p.rho_w; P.rho_sw = p.rho_sw; P.rho_i = p.rho_i; P.g = p.g; P.nu = p.nu; P.n = p.n;
    P.omega = p.omega; P.L = p.L; P.br = p.br; P.lr = p.lr; P.ct = p.ct; P.cw = p.cw; P.p_atm = p.p_atm;
    P.b_min = p.b_min; P.b_max = p.b_max; P.e_v = p.e_v;
    P.n_exp = canon(p.n);
    P.n1_exp = canon(p.n - 1.0);
    P.inv_n_exp = canon(1.0 / p.n);
    P.with_sensible = !(p.ct == 0.0 || p.cw == 0.0);  // src/simulation.jl:132
    P.with_open_by_sliding = !(p.br == 0.0);           // src/simulation.jl:136
    P.kfs = c->m.kfs;
    P.gap_scheme = c->m.gap_scheme;
    P.sliding_law = c->m.sliding_law;
    P.C = c->m.sliding_C;
    P.v0 = c->m.sliding_v0;
    P.Cx = c->slide_cx;
    P.Cy = c->slide_cy;
}

size_t field_count(const shk_ctx* c, int f) {
    if (f < 0 || f >= SHK_NUM_FIELDS) return 0;
    if (f <= SHK_MASK) return c->n;
    if (f <= SHK_VALID_X) return c->nxf;
    return c->nyf;
}

int check_cfg(shk_ctx* ctx, const shk_solver_cfg* cfg) {
    if (cfg->precond != SHK_PRECOND_JACOBI && cfg->precond != SHK_PRECOND_CHEBYSHEV &&
        cfg->precond != SHK_PRECOND_MULTIGRID)
        return fail(ctx, SHK_ERR_ARG, "unknown preconditioner");
    if (cfg->precond == SHK_PRECOND_CHEBYSHEV && (cfg->cheb_degree < 1 || cfg->cheb_degree > MAX_CHEB))
        return fail(ctx, SHK_ERR_ARG, "ChebyshevPreconditioner: degree must be >= 1 (and <= 32)");
    if (cfg->precond == SHK_PRECOND_CHEBYSHEV && (cfg->cheb_nsteps < 2 || cfg->cheb_nsteps > MAX_LANCZOS))
        return fail(ctx, SHK_ERR_ARG, "ChebyshevPreconditioner: nsteps_estimate must be >= 2 (and <= 64)");
    if (cfg->form != SHK_FORM_FP64 && cfg->form != SHK_FORM_FP32_REFINE && cfg->form != SHK_FORM_LEARNED)
        return fail(ctx, SHK_ERR_ARG, "unknown solver form");
    if (cfg->check_every < 1) return fail(ctx, SHK_ERR_ARG, "check_every must be >= 1");
    if (cfg->cheb_bounds != SHK_CHEB_BOUNDS_GERSHGORIN && cfg->cheb_bounds != SHK_CHEB_BOUNDS_REFERENCE)
        return fail(ctx, SHK_ERR_ARG, "unknown Chebyshev bound mode");
    if (!(cfg->atol >= 0) || !(cfg->rtol >= 0)) return fail(ctx, SHK_ERR_ARG, "atol/rtol must be >= 0");
    if (cfg->flags & ~SHK_CFG_PICARD_ATOL) return fail(ctx, SHK_ERR_ARG, "unknown solver flag bits");
    return SHK_OK;
}

// Upload the host-side scalars the next solve starts from.
// The device writes the reset fields itself (no host round trip); the host
// mirror takes the same values and is refreshed whole by the next read.
int reset_scalars(shk_ctx* ctx, double atol, double rtol, int itmax) {
    Scalars* h = ctx->sc_host;
    h->done = 0; h->solved = 0; h->zero_curv = 0; h->tired = 0; h->iter = 0;
    h->atol = atol; h->rtol = rtol; h->itmax = itmax;
    h->counter = 0;
    h->gpend = 0;
    h->done_defer = 0;
    h->breakdown = 0;
    launch::reset_scalars(ctx->sc, atol, rtol, itmax, ctx->st);
    return SHK_OK;
}

int read_scalars(shk_ctx* ctx) {
    HIPCHK(hipMemcpyAsync(ctx->sc_host, ctx->sc, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

// The statistics of a finished Krylov solve from the device scalars (read
// last by the solve loop).  cheb: the Chebyshev preconditioner was in use.
shk_solve_stats stats_of(const shk_ctx* ctx, bool cheb) {
    const Scalars* h = ctx->sc_host;
    shk_solve_stats st;
    memset(&st, 0, sizeof(st));
    st.niter = h->iter;
    st.solved = h->solved;
    st.zero_curvature = h->zero_curv;
    st.tired = h->tired;
    st.bounds_warning = cheb ? h->bounds_warning : 0;
    st.rnorm0 = h->rNorm0;
    st.rnorm = h->rNorm;
    st.lambda_min = h->lmin;
    st.lambda_max = h->lmax;
    st.breakdown = h->breakdown;
    st.breakdown_value = h->breakdown ? h->breakdown_value : 0.0;
    st.lz_lambda_min = cheb ? h->lz_lmin : NAN;
    st.lz_lambda_max = cheb ? h->lz_lmax : NAN;
    st.gershgorin = (cheb && ctx->cfg.cheb_bounds == SHK_CHEB_BOUNDS_GERSHGORIN) ? h->gersh : NAN;
    st.true_residual = NAN;
    return st;
}

void log_solve(shk_ctx* ctx, const shk_solve_stats& st) {
    if (ctx->solve_log.size() >= kLogCap) ctx->solve_log.erase(ctx->solve_log.begin());
    ctx->solve_log.push_back(st);
}

// Record a finished solve (stats out, log) and turn a breakdown into the
// loud failure the reference's Krylov loop gives (sqrt of a negative r'z).
int conclude(shk_ctx* ctx, const shk_solve_stats& st, shk_solve_stats* out) {
    if (out) *out = st;
    log_solve(ctx, st);
    char msg[256];
    if (st.breakdown == 1) {
        snprintf(msg, sizeof msg,
                 "CG breakdown at iteration %d: r'z = %.17g (the preconditioner or the operator is not positive "
                 "definite); h left unchanged", st.niter, st.breakdown_value);
        return fail(ctx, SHK_ERR_NUMERIC, msg);
    }
    if (st.breakdown == 2) {
        snprintf(msg, sizeof msg, "CG breakdown at iteration %d: p'Ap = %.17g is not finite; h left unchanged",
                 st.niter, st.breakdown_value);
        return fail(ctx, SHK_ERR_NUMERIC, msg);
    }
    if (st.breakdown == 3) {
        snprintf(msg, sizeof msg, "the solve's head has a non-finite residual (||b - A h|| / ||b|| = %g)",
                 st.true_residual);
        return fail(ctx, SHK_ERR_NUMERIC, msg);
    }
    return SHK_OK;
}

// Arrays of one linear solve: the stencil coefficients (aP is the diagonal
// in the FP64 form and the flux-form reaction c in the FP32 form), the Jacobi
// diagonal dg, the right-hand side and the CG / Chebyshev workspace.
template <class T>
struct Work {
    const T *aP, *aE, *aN, *dg, *b;
    T *x, *r, *z, *pA, *pB, *Ap, *s, *cA, *cB, *r2;
};

// Chebyshev application z = P^-1 r (passes 1..degree-1), finishing gamma.
template <class T>
void cheb_apply(shk_ctx* c, const Work<T>& w, int degree, int mode) {
    const Geom& G = c->G;
    const double nb = (double)c->n * sizeof(T);
    if (degree == 1) {
        Prof pr(c, P_CHEB, 3 * nb);
        launch::cheb_degree1<T>(G, w.dg, w.r, w.z, mode, c->partials, c->sc, c->st);
        return;
    }
    const T* s_in = nullptr;
    const T* pc_in = nullptr;
    for (int k = 1; k < degree; ++k) {
        T* s_out = w.s;
        T* pc_out = (k % 2) ? w.cA : w.cB;
        const bool last = k == degree - 1;
        // algorithmic bytes (each array once): pass 1 reads r, aP, aE, aN and writes s, pc, y
        // (y only when last); later passes read aP, aE, aN, pc, s, y and write s, pc, y (the
        // last reads r as well and writes y only)
        double bytes = k == 1 ? (last ? 5 : 7) * nb : (last ? 8 : 9) * nb;
        Prof pr(c, P_CHEB, bytes);
        launch::cheb_pass<T>(G, w.aP, w.aE, w.aN, w.dg, w.r, s_in, pc_in, s_out, pc_out, w.z, k, degree, mode,
                             c->partials, c->sc, c->st);
        s_in = s_out;
        pc_in = pc_out;
    }
}

// Lanczos bound estimate + coefficients (Chebyshev only); with the
// Gershgorin mode the bound of D^-1 A first.
template <class T>
void cheb_refresh(shk_ctx* c, const Work<T>& w, int nsteps, int degree) {
    const Geom& G = c->G;
    const double nb = (double)c->n * sizeof(T);
    const int mode = c->cfg.cheb_bounds;
    if (mode == SHK_CHEB_BOUNDS_GERSHGORIN) {
        Prof pr(c, P_CB, 3 * nb);
        launch::gershgorin<T>(G, w.dg, w.aE, w.aN, c->partials, c->sc, c->st);
    }
    const int sym = mode == SHK_CHEB_BOUNDS_GERSHGORIN;
    {
        Prof pr(c, P_LZ, (sym ? 4 : 3) * nb);
        launch::lanczos_init<T>(G, w.b, w.dg, sym, w.r, w.pA, c->partials, c->sc, c->st);
    }
    T* p_old = w.pA;
    T* p_new = w.pB;
    for (int k = 0; k < nsteps; ++k) {
        {
            Prof pr(c, P_LZ, 7 * nb);
            launch::lanczos_step_a<T>(G, w.aP, w.aE, w.aN, w.dg, sym, w.r, p_old, p_new, w.Ap, k, c->partials,
                                      c->sc, c->st);
        }
        if (k < nsteps - 1) {
            Prof pr(c, P_LZ, (sym ? 4 : 3) * nb);
            launch::lanczos_step_b<T>(G, w.Ap, w.dg, sym, w.r, k, c->partials, c->sc, c->st);
        }
        T* t = p_old;
        p_old = p_new;
        p_new = t;
    }
    Prof pr(c, P_CB, 0);
    launch::cheb_bounds(nsteps, degree, 1, mode, c->sc, c->st);
}

// Multigrid hierarchy for the FP32 flux-form system in w (level 0 shares its
// arrays); coarsens 2x2 until at most 64 cells remain (at least once), then
// inverts the coarsest operator densely.
// Whether the finest level takes BF16 coefficients, and their arrays: decided
// before the assembly, which can write them (launch::AsmLow).
int mg_prepare_fine(shk_ctx* ctx) {
    const int nx0 = ctx->g.nx, ny0 = ctx->g.ny;
    ctx->mg_bf16 = ctx->mg_rb && launch::mg_bf16_ok(nx0, ny0) && !getenv("SHK_MG_FP32_FINE");
    if (ctx->mg_bf16 && !ctx->mg_c16) {
        const size_t n0 = (size_t)nx0 * ny0;
        uint16_t** ps[] = {&ctx->mg_c16, &ctx->mg_e16, &ctx->mg_n16, &ctx->mg_D16};
        for (auto pp : ps) HIPCHK(hipMalloc((void**)pp, n0 * sizeof(uint16_t) + 256));
    }
    return SHK_OK;
}

// st waits for the coarsest inverse formed on the side stream, if one is due.
int mg_join_inverse(shk_ctx* ctx) {
    if (!ctx->inv_pending) return SHK_OK;
    HIPCHK(hipStreamWaitEvent(ctx->st, ctx->inv_done, 0));
    ctx->inv_pending = false;
    return SHK_OK;
}

// fine_ready: the finest level's BF16 coefficients were written already (by
// the assembly, from the FP32 values w holds).
int mg_setup(shk_ctx* ctx, const Work<float>& w, bool fine_ready = false) {
    const int nx0 = ctx->g.nx, ny0 = ctx->g.ny;
    if (ctx->mg.empty()) {
        shk_ctx::MGLevel L0{nx0, ny0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, false};
        HIPCHK(hipMalloc((void**)&L0.xa, (size_t)nx0 * ny0 * sizeof(float)));
        ctx->mg.push_back(L0);
        int nx = nx0, ny = ny0;
        while (((long)nx * ny > 64 || ctx->mg.size() == 1) && (long)nx * ny > 1) {
            nx = (nx + 1) / 2;
            ny = (ny + 1) / 2;
            shk_ctx::MGLevel L{nx, ny, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, true};
            const size_t n = (size_t)nx * ny * sizeof(float);
            float** ps[] = {&L.c, &L.e, &L.n, &L.D, &L.b, &L.xa, &L.xb};
            for (auto pp : ps) HIPCHK(hipMalloc((void**)pp, n));
            ctx->mg.push_back(L);
        }
        ctx->mg_tail_first = ctx->mg.size() - 1;
        long tail_cells_max = MG_TAIL_CELLS;  // SHK_MG_TAIL: A/B override, at most MG_TAIL_CELLS
        if (const char* env = getenv("SHK_MG_TAIL")) tail_cells_max = std::min(atol(env), MG_TAIL_CELLS);
        for (size_t l = 1; l < ctx->mg.size(); ++l)
            if ((long)ctx->mg[l].nx * ctx->mg[l].ny <= tail_cells_max) {
                ctx->mg_tail_first = l;
                break;
            }
        // the tail keeps 6 floats per cell of all its levels and the coarsest
        // inverse in LDS: stay within what one workgroup may hold (160 KB on
        // gfx950), less the kernel's 4 KB of static LDS (degenerate shapes
        // such as n x 1 halve only once per level)
        if (const char* e = getenv("SHK_MG_LIN_FROM")) ctx->mg_lin_from = atoi(e);  // negative: none
        if (const char* e = getenv("SHK_MG_NU")) ctx->mg_nu = atoi(e) == 1 ? 1 : 2;
        if (const char* e = getenv("SHK_MG_NU0")) ctx->mg_nu0 = atoi(e) == 1 ? 1 : 2;
        for (size_t l = 0; l + 1 < ctx->mg.size(); ++l)
            ctx->mg[l].lin = ctx->mg_rb && ctx->mg_lin_from >= 0 && (int)l >= ctx->mg_lin_from;
        auto tail_cells = [&](size_t f) {
            size_t c = 0;
            for (size_t l = f; l < ctx->mg.size(); ++l) c += (size_t)ctx->mg[l].nx * ctx->mg[l].ny;
            return c;
        };
        // the linear restriction's staging rows in the tail (coarse width x fine height)
        auto tail_h = [&](size_t f) {
            size_t h = 0;
            for (size_t l = f; l + 1 < ctx->mg.size(); ++l)
                if (ctx->mg[l].lin) h = std::max(h, (size_t)ctx->mg[l + 1].nx * ctx->mg[l].ny);
            return h;
        };
        int dev = 0, lds_max = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t ncl = (size_t)ctx->mg.back().nx * ctx->mg.back().ny;
        const size_t budget = lds_max > 4096 ? (size_t)lds_max - 4096 : 0;
        auto tail_lds = [&](size_t f) { return launch::mg_tail_lds(tail_cells(f), ncl, tail_h(f)); };
        while (ctx->mg_tail_first + 1 < ctx->mg.size() && tail_lds(ctx->mg_tail_first) > budget)
            ++ctx->mg_tail_first;
        // above 64 KB the kernel must opt in; without it, stay within 64 KB
        const size_t dflt = 65536 - 4096;
        if (tail_lds(ctx->mg_tail_first) > dflt && launch::mg_tail_reserve(tail_lds(ctx->mg_tail_first)))
            while (ctx->mg_tail_first + 1 < ctx->mg.size() && tail_lds(ctx->mg_tail_first) > dflt)
                ++ctx->mg_tail_first;
        ctx->mg_tail_h = tail_h(ctx->mg_tail_first);
        if (ctx->mg.size() - ctx->mg_tail_first > (size_t)MG_MAX_TAIL)
            return fail(ctx, SHK_ERR_STATE, "multigrid hierarchy deeper than the tail kernel holds");
        const size_t nc = (size_t)ctx->mg.back().nx * ctx->mg.back().ny;
        // the tail levels' coefficients and the inverse are written straight
        // into the block the tail kernel copies to LDS
        const size_t tp = launch::mg_tail_stride(tail_cells(ctx->mg_tail_first));
        ctx->mg_stride = (int)tp;
        HIPCHK(hipMalloc((void**)&ctx->mg_pack, (4 * tp + nc * nc + 4) * sizeof(float) + 256));
        HIPCHK(hipMemsetAsync(ctx->mg_pack, 0, (4 * tp + nc * nc + 4) * sizeof(float), ctx->st));
        size_t o = 0;
        for (size_t l = ctx->mg_tail_first; l < ctx->mg.size(); ++l) {
            auto& L = ctx->mg[l];
            float** ps[] = {&L.c, &L.e, &L.n, &L.D};
            for (int a = 0; a < 4; ++a) {
                HIPCHK(hipFree(*ps[a]));
                *ps[a] = ctx->mg_pack + a * tp + o;
            }
            L.packed = true;
            o += (size_t)L.nx * L.ny;
        }
        ctx->mg_inv = ctx->mg_pack + 4 * tp;
        if (getenv("SHK_MG_JACOBI")) {
            ctx->mg_omega = 0.95f;
            ctx->mg_scale = ctx->mg_scale_par = 1.5f;
        }
        if (const char* e = getenv("SHK_MG_OMEGA")) ctx->mg_omega = (float)atof(e);
        if (const char* e = getenv("SHK_MG_SCALE")) ctx->mg_scale = (float)atof(e);
        if (const char* e = getenv("SHK_MG_SCALE_PARABOLIC")) ctx->mg_scale_par = (float)atof(e);
        if (const char* e = getenv("SHK_MG_SCALE_LIN")) ctx->mg_scale_lin = (float)atof(e);
        if (const char* e = getenv("SHK_MG_SCALE_LIN_PARABOLIC")) ctx->mg_scale_lin_par = (float)atof(e);
    }
    auto& L0 = ctx->mg[0];
    L0.c = (float*)w.aP;
    L0.e = (float*)w.aE;
    L0.n = (float*)w.aN;
    L0.D = (float*)w.dg;
    if (!fine_ready) {
        int rc = mg_prepare_fine(ctx);
        if (rc) return rc;
        if (ctx->mg_bf16)
            launch::to_bf16x4((size_t)L0.nx * L0.ny, L0.c, L0.e, L0.n, L0.D, ctx->mg_c16, ctx->mg_e16, ctx->mg_n16,
                              ctx->mg_D16, ctx->st);
    }
    // the coarsening rewrites what an inverse still running on the side stream reads
    if (int rc = mg_join_inverse(ctx)) return rc;
    for (size_t l = 1; l < ctx->mg.size(); ++l) {
        auto& f = ctx->mg[l - 1];
        auto& c = ctx->mg[l];
        if (f.lin)
            launch::mg_coarsen_lin(f.nx, f.ny, c.nx, c.ny, f.c, f.e, f.n, c.c, c.e, c.n, ctx->st);
        else
            launch::mg_coarsen(f.nx, f.ny, c.nx, c.ny, f.c, f.e, f.n, c.c, c.e, c.n, ctx->st);
        launch::mg_diag(c.nx, c.ny, c.c, c.e, c.n, c.D, ctx->st);
    }
    auto& Lc = ctx->mg.back();
    if (ctx->prof || !ctx->st_side) {
        Prof pr(ctx, P_MGS, 0);
        launch::mg_dense_inverse(Lc.nx, Lc.ny, Lc.c, Lc.e, Lc.n, ctx->mg_inv, ctx->st);
        return SHK_OK;
    }
    // on the side stream, after the coarsest level's coefficients; the next
    // tail waits for it (mg_join_inverse)
    HIPCHK(hipEventRecord(ctx->inv_fork, ctx->st));
    HIPCHK(hipStreamWaitEvent(ctx->st_side, ctx->inv_fork, 0));
    launch::mg_dense_inverse(Lc.nx, Lc.ny, Lc.c, Lc.e, Lc.n, ctx->mg_inv, ctx->st_side);
    HIPCHK(hipEventRecord(ctx->inv_done, ctx->st_side));
    ctx->inv_pending = true;
    return SHK_OK;
}


// z = V(r): one V-cycle (one damped-Jacobi sweep down and up on every level,
// dense solve on the coarsest), finishing gamma = r'z (mode 0 set-up, 1
// iteration; 2 leaves the partials pending for the next CG matvec; -1 none).
// With T = double the finest level reads the FP64 residual and writes the FP64
// z that FP64 CG uses; every level computes in FP32.  Levels above
// MG_FLAT_CELLS cells run marching kernels, the middle levels single-phase
// kernels, the levels below MG_TAIL_CELLS cells one workgroup.
// zf: where given (FP64 CG with the deferred finish, BF16 finest level), the
// result goes there in FP32, the precision the V-cycle computes it in; the
// CG matvec widens it as it loads it, to the values an FP64 z would hold.
// upd: where given (the same solve, when mg_down_fuses_update), the finest
// down-leg also takes the CG update, reading r and writing r - alpha Ap to
// upd->b_out, which the rest of the V-cycle reads as its right-hand side.
template <class T>
void mg_apply(shk_ctx* ctx, const T* r, T* z, int mode, float* zf = nullptr, const launch::CGUpd* upd = nullptr) {
    auto& mg = ctx->mg;
    const float scale_agg = ctx->system_parabolic ? ctx->mg_scale_par : ctx->mg_scale;
    const float scale_lin = ctx->system_parabolic ? ctx->mg_scale_lin_par : ctx->mg_scale_lin;
    const size_t L = mg.size() - 1;
    const size_t first = ctx->mg_tail_first;  // >= 1
    // bytes per cell of the finest level: its rhs (T), its c, aE, aN, D (BF16 or
    // FP32) and its result (zf: FP32, else T); the coarser levels are all FP32
    const double rb0 = sizeof(T), cb0 = ctx->mg_bf16 ? 8.0 : 16.0, zb0 = zf ? 4.0 : (double)sizeof(T);
    for (size_t l = 0; l < first; ++l) {
        const auto& f = mg[l];
        // reads b, c, aE, aN, D; writes x and the coarse rhs
        const double cells = (double)f.nx * f.ny, coarse = (double)mg[l + 1].nx * mg[l + 1].ny;
        // (with upd also Ap in and r out, FP64)
        Prof pr(ctx, l == 0 ? P_MGD0 : P_MGD,
                l == 0 ? (rb0 + cb0 + 4.0 + (upd ? 16.0 : 0.0)) * cells + 4.0 * coarse
                       : 4.0 * (6.0 * cells + coarse));
        const auto& cl = mg[l + 1];
        bool fused = false;
        if constexpr (sizeof(T) == 8) {
            if (l == 0 && upd) {
                launch::mg_down_bf16_upd(f.nx, f.ny, mg[1].nx, ctx->mg_omega, ctx->mg_c16, ctx->mg_e16, ctx->mg_n16,
                                         ctx->mg_D16, r, f.xa, mg[1].b, *upd, ctx->mg_nu0, ctx->st);
                fused = true;
            }
        }
        if (fused)
            ;
        else if (l == 0 && ctx->mg_bf16)
            launch::mg_down_bf16<T>(f.nx, f.ny, mg[1].nx, ctx->mg_omega, ctx->mg_c16, ctx->mg_e16, ctx->mg_n16,
                                    ctx->mg_D16, r, f.xa, mg[1].b, ctx->sc, ctx->mg_nu0, ctx->st);
        else if (l == 0)
            launch::mg_down<T>(f.nx, f.ny, mg[1].nx, mg[1].ny, ctx->mg_omega, f.c, f.e, f.n, f.D, r, f.xa, mg[1].b,
                               ctx->sc, ctx->mg_rb, ctx->st);
        else if (f.lin && launch::mg_down_lin<float, float>(f.nx, f.ny, cl.nx, ctx->mg_omega, f.c, f.e, f.n, f.D,
                                                                f.b, f.xa, cl.b, ctx->sc, ctx->mg_nu, ctx->st))
            continue;
        else
            launch::mg_down<float>(f.nx, f.ny, mg[l + 1].nx, mg[l + 1].ny, ctx->mg_omega, f.c, f.e, f.n, f.D, f.b,
                                   f.xa, mg[l + 1].b, ctx->sc, ctx->mg_rb, ctx->st);
        if (f.lin) {
            // the linear restriction from the smoothed x the leg stored (its
            // aggregated coarse rhs is overwritten)
            if (l > 0)
                launch::mg_restrict_lin<float, float>(f.nx, f.ny, cl.nx, cl.ny, f.c, f.e, f.n, f.b, f.xa, cl.b,
                                                      ctx->sc, ctx->st);
            else {
                const T* b0 = r;
                if constexpr (sizeof(T) == 8)
                    if (upd) b0 = upd->b_out;
                if (ctx->mg_bf16)
                    launch::mg_restrict_lin<T, uint16_t>(f.nx, f.ny, cl.nx, cl.ny, ctx->mg_c16, ctx->mg_e16,
                                                         ctx->mg_n16, b0, f.xa, cl.b, ctx->sc, ctx->st);
                else
                    launch::mg_restrict_lin<T, float>(f.nx, f.ny, cl.nx, cl.ny, f.c, f.e, f.n, b0, f.xa, cl.b,
                                                      ctx->sc, ctx->st);
            }
        }
    }
    {
        MGTail Tl{};
        Tl.nlev = (int)(L - first + 1);
        Tl.omega = ctx->mg_omega;
        Tl.scale = scale_agg;
        Tl.pack = ctx->mg_pack;
        Tl.stride = ctx->mg_stride;
        Tl.inv = ctx->mg_inv;
        Tl.rb = ctx->mg_rb;
        Tl.scale_lin = scale_lin;
        Tl.hcells = (int)ctx->mg_tail_h;
        for (size_t l = first; l < L; ++l)
            if (mg[l].lin) Tl.lin_mask |= 1 << (l - first);
        double bytes = 0.0;
        for (size_t l = first; l <= L; ++l) {
            const auto& v = mg[l];
            Tl.L[l - first] = MGLevelDev{v.nx, v.ny, v.c, v.e, v.n, v.D, v.b, v.xa, v.xb};
            bytes += 4.0 * 13.0 * v.nx * v.ny;
        }
        Prof pr(ctx, P_MGC, bytes);
        // a failed wait is reported by the entry point's hipGetLastError check
        (void)mg_join_inverse(ctx);
        launch::mg_tail(Tl, ctx->sc, ctx->st);
    }
    const T* r0 = r;  // the finest level's right-hand side on the way up
    if constexpr (sizeof(T) == 8)
        if (upd) r0 = upd->b_out;
    for (size_t l = first; l-- > 0;) {
        const auto& f = mg[l];
        // reads b, c, aE, aN, D, x and the coarse correction; writes the result
        const double cells = (double)f.nx * f.ny, coarse = (double)mg[l + 1].nx * mg[l + 1].ny;
        Prof pr(ctx, l == 0 ? P_MGU0 : P_MGU,
                l == 0 ? (rb0 + cb0 + 4.0 + zb0) * cells + 4.0 * coarse : 4.0 * (7.0 * cells + coarse));
        if (l > 0 && f.lin &&
            launch::mg_up_lin<float, float, float>(f.nx, f.ny, mg[l + 1].nx, ctx->mg_omega, scale_lin, f.c, f.e,
                                                   f.n, f.D, f.b, f.xa, mg[l + 1].xb, f.xb, -1, ctx->partials,
                                                   ctx->sc, ctx->mg_nu, ctx->st))
            continue;
        // the linear prolongation into x first; the leg then adds nothing
        float scale = scale_agg;
        if (f.lin) {
            launch::mg_prolong_lin(f.nx, f.ny, mg[l + 1].nx, scale_lin, f.xa, mg[l + 1].xb, ctx->sc, ctx->st);
            scale = 0.f;
        }
        if (l == 0 && ctx->mg_bf16 && zf)
            launch::mg_up_bf16<T, float>(f.nx, f.ny, mg[1].nx, ctx->mg_omega, scale, ctx->mg_c16, ctx->mg_e16,
                                         ctx->mg_n16, ctx->mg_D16, r0, f.xa, mg[1].xb, zf, mode, ctx->partials,
                                         ctx->sc, ctx->mg_nu0, ctx->st);
        else if (l == 0 && ctx->mg_bf16)
            launch::mg_up_bf16<T, T>(f.nx, f.ny, mg[1].nx, ctx->mg_omega, scale, ctx->mg_c16, ctx->mg_e16,
                                     ctx->mg_n16, ctx->mg_D16, r0, f.xa, mg[1].xb, z, mode, ctx->partials, ctx->sc,
                                     ctx->mg_nu0, ctx->st);
        else if (l == 0)
            launch::mg_up<T>(f.nx, f.ny, mg[1].nx, ctx->mg_omega, scale, f.c, f.e, f.n, f.D, r, f.xa,
                             mg[1].xb, z, mode, ctx->partials, ctx->sc, ctx->mg_rb, ctx->st);
        else
            launch::mg_up<float>(f.nx, f.ny, mg[l + 1].nx, ctx->mg_omega, scale, f.c, f.e, f.n, f.D, f.b,
                                 f.xa, mg[l + 1].xb, f.xb, -1, ctx->partials, ctx->sc, ctx->mg_rb, ctx->st);
    }
}

// Learned preconditioner (FP64 CG only): z = gamma e + P(r - gamma A e), with
// e the network's correction for r, gamma = e'r / e'Ae its energy-minimising
// step and P the fixed preconditioner; then gamma = r'z and the flexible-CG
// beta = -z'Ap / p'Ap.  r - gamma A e goes through w.r2 (unused unless the
// fused Chebyshev kernel runs, which learned batches do not use).
static int nn_precond(shk_ctx* ctx, const Work<double>& w, int precond, int degree) {
    const size_t n = ctx->n;
    {
        Prof pr(ctx, P_NNP, 0);
        launch::nn_rscale(n, ctx->aP, w.r, ctx->partials, ctx->sc, ctx->st);
        launch::nn_features(ctx->g.nx, ctx->g.ny, ctx->W8, ctx->H8, ctx->aP, ctx->aE, ctx->aN, ctx->cP, w.r, ctx->sc,
                            1, ctx->nn_x, ctx->st);
    }
    if (ctx->rec_kind == 1 && ctx->rec_x && ctx->rec_count < ctx->rec_cap) {
        // training pair of this iteration: its input now, its target (the
        // remaining error) once the solve has converged
        const size_t k = (size_t)ctx->rec_count++;
        const size_t img = (size_t)ctx->W8 * ctx->H8;
        HIPCHK(hipMemcpyAsync((char*)ctx->rec_x + k * img * 16, ctx->nn_x, img * 16, hipMemcpyDeviceToDevice,
                              ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->rec_snap + k * n, w.x, n * 8, hipMemcpyDeviceToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->rec_scale + k, &ctx->sc->nn_scale_pc, 8, hipMemcpyDeviceToDevice, ctx->st));
    }
    {
        Prof pr(ctx, P_NNF, 0);
        if (shknn::forward(ctx->net, ctx->nn_x, 1, ctx->H8, ctx->W8))
            return fail(ctx, SHK_ERR_STATE, std::string("network: ") + shknn::unet_error(ctx->net));
    }
    {
        Prof pr(ctx, P_NNP, 0);
        launch::nn_correction(ctx->g.nx, ctx->g.ny, ctx->W8, shknn::output(ctx->net), ctx->aP, ctx->aE, ctx->aN,
                              ctx->sc, 1, ctx->nn_e, ctx->st);
        launch::matvec<double>(ctx->G, ctx->aP, ctx->aE, ctx->aN, ctx->nn_e, ctx->nn_Ae, ctx->st);
        launch::nn_line_search(n, ctx->nn_e, w.r, ctx->nn_Ae, ctx->partials, ctx->sc, 1, ctx->st);
        launch::nn_r1(n, w.r, ctx->nn_Ae, ctx->sc, w.r2, ctx->st);
    }
    if (precond == SHK_PRECOND_MULTIGRID) {
        mg_apply<double>(ctx, w.r2, w.z, -1);
    } else {
        Work<double> w1 = w;
        w1.r = w.r2;
        cheb_apply<double>(ctx, w1, degree, -1);
    }
    Prof pr(ctx, P_NNP, 0);
    launch::nn_combine(n, w.r, ctx->nn_e, w.z, ctx->partials, ctx->sc, ctx->st);
    launch::fcg_beta(n, w.z, w.Ap, ctx->partials, ctx->sc, ctx->st);
    return SHK_OK;
}

// Preconditioned CG on A x = b from x0 (nullptr: zero), Krylov cg! semantics.
// The accumulated correction is left in w.x.  Returns SHK_OK or an error.
//
// nn_every > 0 (FP64 with an attached network, multigrid or Chebyshev):
// flexible CG.  Each batch of nn_every iterations starts with one iteration
// whose preconditioner is the learned one; that iteration and the next take
// the flexible (truncated, one-vector) beta = -z'Ap / p'Ap, which keeps p
// A-orthogonal to the previous direction when the preconditioner changes;
// the others keep Krylov's beta, which equals it for a fixed preconditioner.
// The first batch is plain PCG and sets the reference reduction rho_ref =
// (residual norm at its end) / (at its start).  A learned batch whose own
// reduction is not below nn_stall * rho_ref ends the learned batches: the
// solve goes on as plain PCG from where it is (*nn_stalled = 1).
template <class T>
int pcg(shk_ctx* ctx, const Work<T>& w, const T* x0, int precond, int degree, int check_every, double atol,
        double rtol, int itmax, int nn_every = 0, int* nn_stalled = nullptr, bool started = false) {
    const Geom& G = ctx->G;
    const double nb = (double)ctx->n * sizeof(T);
    const int jacobi = precond == SHK_PRECOND_JACOBI;
    int rc = reset_scalars(ctx, atol, rtol, itmax);
    if (rc) return rc;
    // started: r, x and p_zero hold the start from x0 already (no Jacobi)
    if (!started || jacobi) {
        Prof pr(ctx, P_INIT, (x0 ? 8 : 5) * nb);
        launch::cg_init<T>(G, w.aP, w.aE, w.aN, w.dg, w.b, x0, w.r, w.z, w.x, w.pA, jacobi, ctx->partials, ctx->sc,
                           ctx->st);
    }
    const bool multigrid = precond == SHK_PRECOND_MULTIGRID;
    // FP64 CG with the V-cycle and the deferred finish: z kept in FP32 (the
    // precision the V-cycle produces it in) where the finest level is BF16 and
    // the matvec is the two-column kernel
    const bool zf32 = sizeof(T) == 8 && multigrid && nn_every == 0 && ctx->use_defer && ctx->mg_bf16 &&
                      cg_wide_ok(G);
    float* const zf = zf32 ? ctx->z32 : nullptr;
    if (multigrid) {
        mg_apply<T>(ctx, w.r, w.z, 0, zf);
    } else if (!jacobi) {
        cheb_apply<T>(ctx, w, degree, 0);
    }
    // an even batch returns the p ping-pong buffers to their starting roles,
    // so one captured batch can be replayed
    const bool nn = nn_every > 0;
    if constexpr (sizeof(T) != 8) {
        if (nn) return fail(ctx, SHK_ERR_STATE, "the learned preconditioner runs in FP64 CG only");
    }
    const int every = nn_every + (nn_every & 1);
    const int bmax = nn ? every : check_every + (check_every & 1);
    const bool fused = precond == SHK_PRECOND_CHEBYSHEV && degree >= 2 && degree <= 8 && ctx->use_fused && !nn;
    bool learned_batch = false;
    int rc_nn = SHK_OK;
    // FP64 CG with the V-cycle (no learned preconditioner): the two finishing
    // kernels of an iteration are folded into the matvec and the update
    const bool defer = sizeof(T) == 8 && multigrid && !nn && ctx->use_defer;
    // plain solves size each batch to the iterations the convergence rate so
    // far predicts (1..bmax), so few iterations are queued past convergence.
    // The first batch takes the count of the same Picard iteration of the
    // previous loop (else of the previous solve), less batch_margin: one more
    // batch costs a host round trip, about what one iteration queued past
    // convergence costs (its launches read the done flag and exit, a few
    // microseconds each).  A deferred batch ends with the stop decision the
    // next matvec would take (cg_defer_stop), so it needs no iteration beyond
    // the solve's own.  An odd batch leaves the p (and r) ping-pong buffers in
    // swapped roles: the next batch starts from the other phase.
    const bool adaptive = !nn && ctx->use_adaptive;
    auto size_batch = [&](double need) {
        const int b = need <= 1.0 ? 1 : (need >= bmax ? bmax : (int)ceil(need));
        return b < bmax ? b : bmax;
    };
    int predicted = ctx->last_niter;
    if constexpr (sizeof(T) == 8) {
        const int k = ctx->picard_it;
        if (k > 0 && k <= (int)ctx->niter_by_picard.size() && ctx->niter_by_picard[k - 1] > 0)
            predicted = ctx->niter_by_picard[k - 1];
    }
    int batch = (adaptive && predicted > 0) ? size_batch((double)predicted - ctx->batch_margin) : bmax;
    int phase = 0;  // the ping-pong roles at the start of the next batch
    // ... with the CG update fused into the finest down-leg where that leg can
    // take it (r then alternates between w.r and w.r2; batches are even)
    const bool fuse_upd = defer && zf32 && launch::mg_down_fuses_update(ctx->mg[0].nx, ctx->mg[0].ny) &&
                          getenv("SHK_NO_FUSED_UPDATE") == nullptr;
    const int nba = multigrid ? launch::mg_up_nblocks(ctx->mg[0].nx, ctx->mg[0].ny, ctx->mg_rb,
                                                      ctx->mg_bf16 ? ctx->mg_nu0 : 1)
                              : 0;
    // x steps in pairs: iteration j (from 1) is odd where phase + k is even
    const bool pairx = defer && zf32 && ctx->pair_x;
    auto xs_of = [&](int k) { return pairx ? (((phase + k) & 1) == 0 ? 0 : 2) : 1; };
    auto enqueue_batch = [&]() {
        T* p_old = phase ? w.pB : w.pA;
        T* p_new = phase ? w.pA : w.pB;
        T* r_cur = phase ? w.r2 : w.r;
        T* r_nxt = phase ? w.r : w.r2;
        for (int k = 0; k < batch; ++k) {
            if constexpr (sizeof(T) == 8) {
                if (defer) {
                    int nbb;
                    {
                        // reads aP, aE, aN, z (FP32 or FP64), p_old, x; writes p_new, Ap, x
                        // (x += alpha p_old: the previous iteration's step); paired
                        // steps: no x (odd), or also p_new read (even)
                        const int xs = xs_of(k);
                        Prof pr(ctx, P_MV, (zf32 ? 8.5 : 9.0) * nb + (xs == 0 ? -2.0 : xs == 2 ? 1.0 : 0.0) * nb);
                        nbb = zf32 ? launch::cg_matvec_deferred(G, w.aP, w.aE, w.aN, (const float*)zf, p_old, p_new,
                                                                w.Ap, ctx->partials2, ctx->sc, ctx->partials, nba,
                                                                w.x, ctx->st, xs)
                                   : launch::cg_matvec_deferred(G, w.aP, w.aE, w.aN, w.z, p_old, p_new, w.Ap,
                                                                ctx->partials2, ctx->sc, ctx->partials, nba, w.x,
                                                                ctx->st);
                    }
                    if (fuse_upd) {
                        // the update in the finest down-leg: r_cur in, r_cur - alpha Ap to r_nxt
                        const launch::CGUpd U{w.Ap, r_nxt, ctx->sc, ctx->partials, nba, ctx->partials2, nbb};
                        mg_apply<T>(ctx, r_cur, w.z, 2, zf, &U);
                        T* tr = r_cur;
                        r_cur = r_nxt;
                        r_nxt = tr;
                    } else {
                        {
                            Prof pr(ctx, P_UPD, 3 * nb);  // reads r, Ap; writes r
                            launch::cg_update_deferred(G, w.Ap, w.r, ctx->sc, ctx->partials, nba, ctx->partials2,
                                                       nbb, ctx->st);
                        }
                        mg_apply<T>(ctx, w.r, w.z, 2, zf);
                    }
                    T* t = p_old;
                    p_old = p_new;
                    p_new = t;
                    continue;
                }
            }
            {
                Prof pr(ctx, P_MV, 7 * nb);  // reads aP, aE, aN, z, p_old; writes p_new, Ap
                launch::cg_matvec<T>(G, w.aP, w.aE, w.aN, w.z, p_old, p_new, w.Ap, ctx->partials, ctx->sc, ctx->st);
            }
            if (fused) {
                // reads aP, aE, aN, p, Ap, x, r_old; writes x, r_new, z
                Prof pr(ctx, P_FUSED, 10 * nb);
                launch::cg_update_cheb<T>(G, degree, w.aP, w.aE, w.aN, w.dg, p_new, w.Ap, r_cur, r_nxt, w.x, w.z,
                                          ctx->partials, ctx->sc, ctx->st);
                T* t = r_cur;
                r_cur = r_nxt;
                r_nxt = t;
            } else {
                {
                    Prof pr(ctx, P_UPD, (jacobi ? 8 : 6) * nb);
                    launch::cg_update<T>(G, w.dg, p_new, w.Ap, w.x, w.r, w.z, jacobi, ctx->partials, ctx->sc,
                                         ctx->st);
                }
                if constexpr (sizeof(T) == 8) {
                    if (learned_batch && k == 0) {
                        if (!rc_nn) rc_nn = nn_precond(ctx, w, precond, degree);
                    } else {
                        if (multigrid) mg_apply<T>(ctx, w.r, w.z, 1);
                        else if (!jacobi) cheb_apply<T>(ctx, w, degree, 1);
                        if (learned_batch && k == 1) {
                            Prof pr(ctx, P_NNP, 0);
                            launch::fcg_beta(ctx->n, w.z, w.Ap, ctx->partials, ctx->sc, ctx->st);
                        }
                    }
                } else {
                    if (multigrid) mg_apply<T>(ctx, w.r, w.z, 1);
                    else if (!jacobi) cheb_apply<T>(ctx, w, degree, 1);
                }
            }
            T* t = p_old;
            p_old = p_new;
            p_new = t;
        }
        if constexpr (sizeof(T) == 8) {
            if (defer) {
                Prof pr(ctx, P_FIN, 0);
                launch::cg_defer_stop(ctx->n, ctx->sc, ctx->partials, nba, p_old, w.x, ctx->st, xs_of(batch),
                                      p_new);
            }
        }
    };
    // one graph per batch kind; the learned batch's graph holds the network's
    // kernels too (they run on ctx->st), so it is keyed on the network and
    // its image
    auto batch_graph = [&](hipGraphExec_t* out) -> int {
        *out = nullptr;
        if (!ctx->use_graphs || ctx->prof || (ctx->rec_kind == 1 && ctx->rec_x)) return SHK_OK;
        // the V-cycle's coarse scale (a kernel argument) differs between
        // elliptic and parabolic systems: part of the key
        const int key[4] = {(int)sizeof(T),
                            precond + (learned_batch ? 16 : 0) + (defer ? 32 : 0) + (ctx->system_parabolic ? 64 : 0) +
                                (fuse_upd ? 128 : 0) + (phase ? 256 : 0),
                            degree * (fused ? 100 : 1), batch};
        for (auto& e : ctx->graphs)
            if (!memcmp(e.key, key, sizeof(key)) && (!learned_batch || e.net == ctx->net)) *out = e.exec;
        if (*out) return SHK_OK;
        shk_ctx::GraphEntry e;
        memcpy(e.key, key, sizeof(key));
        e.net = learned_batch ? ctx->net : nullptr;
        if (int rc = mg_join_inverse(ctx)) return rc;  // no event from outside the capture inside it
        HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
        enqueue_batch();
        HIPCHK(hipStreamEndCapture(ctx->st, &e.graph));
        if (rc_nn) {
            hipGraphDestroy(e.graph);
            return rc_nn;
        }
        HIPCHK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        ctx->graphs.push_back(e);
        *out = e.exec;
        return SHK_OK;
    };
    hipGraphExec_t exec = nullptr;
    rc = batch_graph(&exec);
    if (rc) return rc;
    if (nn) {
        // the residual norm the first learned batch is measured against (and
        // no network run when the start already meets the test)
        rc = read_scalars(ctx);
        if (rc) return rc;
        if (ctx->sc_host->done) return SHK_OK;
    }
    double rn_start = ctx->sc_host->rNorm, rho_ref = 0.0;
    bool plain_done = false;
    long queued = 0;
    for (;;) {
        if (exec) HIPCHK(hipGraphLaunch(exec, ctx->st));
        else enqueue_batch();
        if (rc_nn) return rc_nn;
        queued += batch;
        phase ^= batch & 1;
        if (ctx->prof) prof_collect(ctx);
        rc = read_scalars(ctx);
        if (rc) return rc;
        if (ctx->sc_host->done || ctx->sc_host->done_defer) break;
        if (queued > (long)itmax + batch) break;  // cannot happen: tired sets done
        if (adaptive) {
            const Scalars* h = ctx->sc_host;
            const double prev = queued == batch ? h->rNorm0 : rn_start;
            const double ratio = h->rNorm / prev;
            double need = bmax;
            if (ratio > 0.0 && ratio < 1.0 && h->rNorm > 0.0 && h->eps_stop > 0.0) {
                const double per_it = log(ratio) / batch;  // log of the mean reduction per iteration
                need = ceil(log(h->eps_stop / h->rNorm) / per_it);
            }
            batch = size_batch(need);
            rc = batch_graph(&exec);  // the size or the phase may have changed
            if (rc) return rc;
        }
        if (nn && !plain_done) {
            // the plain batch's reduction is the reference; learned batches follow
            plain_done = true;
            rho_ref = ctx->sc_host->rNorm / rn_start;
            learned_batch = true;
            rc = batch_graph(&exec);
            if (rc) return rc;
        } else if (learned_batch) {
            const double rho = ctx->sc_host->rNorm / rn_start;
            if (!(rho < ctx->nn_stall * rho_ref)) {
                // stalled: plain PCG from here on (the learned iteration was the
                // first of the batch, so no later iteration needs the flexible beta)
                learned_batch = false;
                if (nn_stalled) *nn_stalled = 1;
                rc = batch_graph(&exec);
                if (rc) return rc;
            }
        }
        rn_start = ctx->sc_host->rNorm;
    }
    if constexpr (sizeof(T) == 8) {
        if (pairx) {
            Prof pr(ctx, P_FIN, 0);
            launch::cg_xflush(ctx->n, ctx->sc, w.x, ctx->st);
        }
    }
    ctx->last_niter = ctx->sc_host->iter;
    if constexpr (sizeof(T) == 8) {
        const int k = ctx->picard_it;
        if (k > 0) {
            if ((int)ctx->niter_by_picard.size() < k) ctx->niter_by_picard.resize(k, 0);
            ctx->niter_by_picard[k - 1] = ctx->sc_host->iter;
        }
    }
    return SHK_OK;
}

// FP32-inner / FP64-refinement solve (SHK_FORM_FP32_REFINE).  The FP64
// residual r = b - A h is measured in the Jacobi-weighted norm sqrt(r'D^-1 r);
// each refinement solves A d = r in FP32 (FP32 operator and Chebyshev sweeps,
// FP64 dot products) to inner_rtol and adds d to the FP64 iterate, until the
// FP64 norm meets atol + rtol * initial (the same relative test as the FP64 form).
int solve_ir(shk_ctx* ctx, int parabolic, double dt, double relax_alpha, shk_solve_stats* stats) {
    const Geom& G = ctx->G;
    const size_t n = ctx->n;
    const double nb = (double)n * 8;
    const shk_solver_cfg& cfg = ctx->cfg;
    const int itmax = cfg.itmax > 0 ? cfg.itmax : (int)(2 * n < 2147483647 ? 2 * n : 2147483647);
    {
        Prof pr(ctx, P_ASM, 17 * nb);
        launch::assemble<double>(G, ctx->P, ctx->F, parabolic, dt, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->st,
                                 ctx->cP);
    }
    ctx->have_system = true;
    ctx->system_parabolic = parabolic;
    ctx->system_dt = dt;
    {
        // FP32 flux-form system: c (in aP32), aE, aN, and the Jacobi diagonal
        Prof pr(ctx, P_IR, 5 * 12.0 * n);
        launch::to_f32(n, ctx->cP, ctx->aP32, ctx->st);
        launch::to_f32(n, ctx->aE, ctx->aE32, ctx->st);
        launch::to_f32(n, ctx->aN, ctx->aN32, ctx->st);
        launch::to_f32(n, ctx->aP, ctx->dg32, ctx->st);
        launch::to_f32(n, ctx->rhs, ctx->b32, ctx->st);
    }
    // b32 holds the FP32 rhs for the bound estimate, then each FP32 residual:
    // it is the right-hand side of every inner solve
    float* r_in = ctx->b32;
    Work<float> wi{ctx->aP32, ctx->aE32, ctx->aN32, ctx->dg32, r_in, ctx->x32, ctx->r32, ctx->z32, ctx->pA32,
                   ctx->pB32, ctx->Ap32, ctx->s32, ctx->cA32, ctx->cB32, ctx->r2_32};
    const int cheb = cfg.precond == SHK_PRECOND_CHEBYSHEV;
    // bounds of the FP32 operator, estimated from the FP32 rhs (scale-free)
    if (cheb) cheb_refresh<float>(ctx, wi, cfg.cheb_nsteps, cfg.cheb_degree);
    if (cfg.precond == SHK_PRECOND_MULTIGRID) {
        int rc0 = mg_setup(ctx, wi);
        if (rc0) return rc0;
    }
    double* x = ctx->h_alt;  // the FP64 iterate, started from h
    double* h_prev = ctx->F.f[SHK_H];
    HIPCHK(hipMemcpyAsync(x, h_prev, n * 8, hipMemcpyDeviceToDevice, ctx->st));
    int rc = reset_scalars(ctx, 0.0, 0.0, itmax);
    if (rc) return rc;
    ctx->sc_host->ir_atol = std::max(cfg.atol, ctx->atol_floor);
    ctx->sc_host->ir_rtol = cfg.rtol;
    HIPCHK(hipMemcpyAsync(&ctx->sc->ir_atol, &ctx->sc_host->ir_atol, 2 * sizeof(double), hipMemcpyHostToDevice,
                          ctx->st));
    int refinements = 0;
    long inner = 0;
    double rnorm0 = 0.0, rnorm = 0.0, best = INFINITY;
    int solved = 0, zc = 0, tired = 1, stall = 0;
    constexpr int kMaxRefinements = 64;
    for (int k = 0; k < kMaxRefinements; ++k) {
        {
            Prof pr(ctx, P_IR, 8.5 * nb);
            launch::ir_residual(G, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, x, r_in, k == 0, ctx->partials, ctx->sc,
                                ctx->st);
        }
        rc = read_scalars(ctx);
        if (rc) return rc;
        if (k == 0) rnorm0 = ctx->sc_host->ir_rnorm0;
        rnorm = ctx->sc_host->ir_rnorm;
        if (ctx->sc_host->ir_done) { solved = 1; tired = 0; break; }
        // the precision floor: three refinements in a row that do not lower
        // the FP64 residual by 10% end the loop (neither solved nor tired; a
        // target below what FP64 can represent, as after Picard convergence)
        if (rnorm < 0.9 * best) {
            best = rnorm;
            stall = 0;
        } else if (++stall >= 3) {
            tired = 0;
            break;
        }
        rc = pcg<float>(ctx, wi, nullptr, cfg.precond, cfg.cheb_degree, cfg.check_every, 0.0, cfg.inner_rtol,
                        itmax);
        if (rc) return rc;
        inner += ctx->sc_host->iter;
        zc |= ctx->sc_host->zero_curv;
        if (ctx->sc_host->breakdown) {
            // an inner solve broke down: h is left as it was
            shk_solve_stats sst = stats_of(ctx, cheb);
            sst.niter = (int32_t)inner;
            sst.refinements = refinements;
            sst.rnorm0 = rnorm0;
            sst.rnorm = rnorm;
            if (ctx->prof) prof_collect(ctx);
            return conclude(ctx, sst, stats);
        }
        {
            Prof pr(ctx, P_IR, 20.0 * n);
            launch::ir_correct(n, x, ctx->x32, ctx->st);
        }
        ++refinements;
        if (ctx->sc_host->iter == 0) {
            // the inner solve met Krylov's stopping test at its start (its
            // absolute test rNorm + 1 <= 1 on a residual at the FP64 floor):
            // no refinement can lower it further
            solved = ctx->sc_host->solved;
            tired = 0;
            break;
        }
    }
    shk_solve_stats sst = stats_of(ctx, cheb);
    sst.niter = (int32_t)inner;
    sst.solved = solved;
    sst.zero_curvature = zc;
    sst.tired = tired;
    sst.refinements = refinements;
    sst.rnorm0 = rnorm0;
    sst.rnorm = rnorm;
    {
        Prof pr(ctx, P_FIN, 6 * nb);
        launch::residual_norms(G, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, x, nullptr, ctx->partials, &ctx->sc->tres,
                               ctx->st);
    }
    rc = read_scalars(ctx);
    if (rc) return rc;
    sst.true_residual = ctx->sc_host->tres_b > 0.0 ? ctx->sc_host->tres / ctx->sc_host->tres_b : ctx->sc_host->tres;
    if (!std::isfinite(sst.true_residual)) {
        sst.breakdown = 3;
        if (ctx->prof) prof_collect(ctx);
        return conclude(ctx, sst, stats);
    }
    if (relax_alpha > 0.0) launch::relax(n, x, h_prev, relax_alpha, ctx->st);
    ctx->F.f[SHK_H] = x;
    ctx->h_alt = h_prev;
    if (ctx->prof) prof_collect(ctx);
    return conclude(ctx, sst, stats);
}

int solve(shk_ctx* ctx, int parabolic, double dt, double relax, shk_solve_stats* stats) {
    {
        // the stopping test refers to the CG start's own residual unless the
        // learned warm start below sets the reference to r(h)
        ctx->sc_host->rnorm0_ref = 0.0;
        HIPCHK(hipMemcpyAsync(&ctx->sc->rnorm0_ref, &ctx->sc_host->rnorm0_ref, sizeof(double),
                              hipMemcpyHostToDevice, ctx->st));
    }
    if (ctx->cfg.form == SHK_FORM_FP32_REFINE) return solve_ir(ctx, parabolic, dt, relax, stats);
    const Geom& G = ctx->G;
    const double nb = (double)ctx->n * 8;
    const shk_solver_cfg& cfg = ctx->cfg;
    const int itmax = cfg.itmax > 0 ? cfg.itmax : (int)(2 * ctx->n < 2147483647 ? 2 * ctx->n : 2147483647);
    const bool multigrid = cfg.precond == SHK_PRECOND_MULTIGRID;
    const bool learned = cfg.form == SHK_FORM_LEARNED;
    if (learned && !ctx->net)
        return fail(ctx, SHK_ERR_STATE, "form SHK_FORM_LEARNED needs a network: call shk_attach_unet first");
    const bool record = ctx->rec_kind == 0 && ctx->rec_x && ctx->rec_count < ctx->rec_cap;
    const bool cheb_pc = cfg.precond == SHK_PRECOND_CHEBYSHEV;
    if (ctx->rec_kind == 1 && ctx->rec_x && !(learned && ctx->nn_every > 0 && (multigrid || cheb_pc)))
        return fail(ctx, SHK_ERR_STATE,
                    "recording CG-iteration pairs needs SHK_FORM_LEARNED with a learned preconditioner period "
                    "and the multigrid or Chebyshev preconditioner");
    const int rec_first = ctx->rec_count;
    const bool need_c = multigrid || learned || record;
    // FP64 CG with the V-cycle: the assembly also writes the FP32 flux form the
    // hierarchy is built from and the finest level's BF16 coefficients
    launch::AsmLow lo;
    // the CG start from h, formed by the assembly (nothing runs between the
    // two that reads or writes r, x or p)
    const bool asm_start = multigrid && !learned && !record && ctx->asm_start;
    launch::AsmInit ini;
    if (asm_start) ini = launch::AsmInit{ctx->r, ctx->x, ctx->pA};
    if (multigrid) {
        int rc0 = mg_prepare_fine(ctx);
        if (rc0) return rc0;
        lo.c = ctx->aP32;
        lo.e = ctx->aE32;
        lo.n = ctx->aN32;
        lo.d = ctx->dg32;
        if (ctx->mg_bf16) {
            lo.c16 = ctx->mg_c16;
            lo.e16 = ctx->mg_e16;
            lo.n16 = ctx->mg_n16;
            lo.d16 = ctx->mg_D16;
        }
    }
    {
        Prof pr(ctx, P_ASM, (need_c ? 18 : 17) * nb + (multigrid ? (ctx->mg_bf16 ? 3.0 : 2.0) * nb : 0.0) +
                                (asm_start ? 3.0 * nb : 0.0));
        launch::assemble<double>(G, ctx->P, ctx->F, parabolic, dt, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->st,
                                 need_c ? ctx->cP : nullptr, lo, ini);
    }
    ctx->have_system = true;
    ctx->system_parabolic = parabolic;
    ctx->system_dt = dt;
    Work<double> w{ctx->aP, ctx->aE, ctx->aN, ctx->aP, ctx->rhs, ctx->x, ctx->r, ctx->z, ctx->pA, ctx->pB,
                   ctx->Ap, ctx->s, ctx->cA, ctx->cB, ctx->r2};
    const int cheb = cfg.precond == SHK_PRECOND_CHEBYSHEV;
    if (cheb) cheb_refresh<double>(ctx, w, cfg.cheb_nsteps, cfg.cheb_degree);
    if (multigrid) {
        // FP64 CG preconditioned by the FP32 V-cycle: the hierarchy is built
        // from the FP32 flux form of this system (written by the assembly)
        Work<float> wi{ctx->aP32, ctx->aE32, ctx->aN32, ctx->dg32, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        int rc0 = mg_setup(ctx, wi, true);
        if (rc0) return rc0;
    }
    const double* x0 = ctx->F.f[SHK_H];
    const size_t slot = (size_t)ctx->rec_count * ctx->W8 * ctx->H8;
    if (learned || record) {
        int rc0 = reset_scalars(ctx, std::max(cfg.atol, ctx->atol_floor), cfg.rtol, itmax);
        if (rc0) return rc0;
        void* xin = record ? (void*)((char*)ctx->rec_x + slot * 16) : ctx->nn_x;
        {
            // r(h), its scale, and the network input
            Prof pr(ctx, P_NNW, 0);
            launch::matvec<double>(G, ctx->aP, ctx->aE, ctx->aN, x0, w.Ap, ctx->st);
            launch::nn_resid_scale(ctx->n, ctx->aP, ctx->rhs, w.Ap, w.r, ctx->partials, ctx->sc, ctx->st);
            launch::nn_features(ctx->g.nx, ctx->g.ny, ctx->W8, ctx->H8, ctx->aP, ctx->aE, ctx->aN, ctx->cP, w.r,
                                ctx->sc, 0, xin, ctx->st);
        }
        if (learned) {
            // the preconditioned norm of r(h): the stopping test's reference,
            // so a better start means fewer iterations, not a tighter target
            if (multigrid) mg_apply<double>(ctx, w.r, w.z, 0);
            else if (cheb) cheb_apply<double>(ctx, w, cfg.cheb_degree, 0);
            launch::nn_ref(ctx->sc, (multigrid || cheb) ? 0 : 1, (double)ctx->n, ctx->st);
        }
        if (learned) {
            {
                Prof pr(ctx, P_NNF, 0);
                shknn::set_stream(ctx->net, ctx->st);
                if (shknn::forward(ctx->net, xin, 1, ctx->H8, ctx->W8))
                    return fail(ctx, SHK_ERR_STATE, std::string("network: ") + shknn::unet_error(ctx->net));
            }
            Prof pr(ctx, P_NNW, 0);
            launch::nn_correction(ctx->g.nx, ctx->g.ny, ctx->W8, shknn::output(ctx->net), ctx->aP, ctx->aE, ctx->aN,
                                  ctx->sc, 0, ctx->nn_e, ctx->st);
            launch::matvec<double>(G, ctx->aP, ctx->aE, ctx->aN, ctx->nn_e, ctx->nn_Ae, ctx->st);
            launch::nn_line_search(ctx->n, ctx->nn_e, w.r, ctx->nn_Ae, ctx->partials, ctx->sc, 0, ctx->st);
            launch::nn_warm(ctx->n, x0, ctx->nn_e, ctx->sc, ctx->hw, ctx->st);
            x0 = ctx->hw;
        }
    }
    // the learned preconditioner needs a fixed preconditioner to compose with
    const int nn_every = (learned && (multigrid || cheb)) ? ctx->nn_every : 0;
    int stalled = 0;
    int rc = pcg<double>(ctx, w, x0, cfg.precond, cfg.cheb_degree, cfg.check_every,
                         std::max(cfg.atol, ctx->atol_floor), cfg.rtol, itmax,
                         nn_every, &stalled, asm_start && x0 == ctx->F.f[SHK_H]);
    if (rc) return rc;
    shk_solve_stats sst = stats_of(ctx, cheb);
    sst.fallback = learned ? (ctx->sc_host->nn_fallback | (stalled ? 2 : 0)) : 0;
    if (sst.breakdown) {
        if (ctx->prof) prof_collect(ctx);
        return conclude(ctx, sst, stats);
    }
    // the FP64 residual of the head about to be returned, x + x0; in the usual
    // case (warm start from h) the same kernel writes that head, under-relaxed,
    // into the alternate buffer, which becomes h only when the residual is finite
    double* h_prev = ctx->F.f[SHK_H];
    const bool fused_finish = x0 == h_prev && !(ctx->rec_kind == 1 && ctx->rec_x);
    // ... and, for a Picard loop, the Picard test of that head against h
    const bool fuse_pic = fused_finish && ctx->picard_fuse;
    {
        Prof pr(ctx, P_FIN, (fused_finish ? 7 : 6) * nb);
        launch::residual_norms(G, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->x, x0, ctx->partials, &ctx->sc->tres,
                               ctx->st, fused_finish ? ctx->h_alt : nullptr, relax,
                               fuse_pic ? ctx->partials2 : nullptr, ctx->sc);
    }
    rc = read_scalars(ctx);
    if (rc) return rc;
    ctx->picard_fused = fuse_pic;
    {
        const Scalars* h = ctx->sc_host;
        sst.true_residual = h->tres_b > 0.0 ? h->tres / h->tres_b : h->tres;
        if (!std::isfinite(sst.true_residual)) {
            sst.breakdown = 3;
            if (ctx->prof) prof_collect(ctx);
            return conclude(ctx, sst, stats);
        }
    }
    if (ctx->rec_kind == 1 && ctx->rec_x) {
        // targets of the pairs this solve recorded: sqrt(aP) (x_final - x_k) / s_k
        const size_t img = (size_t)ctx->W8 * ctx->H8;
        for (int k = rec_first; k < ctx->rec_count; ++k)
            launch::nn_target_cg(ctx->g.nx, ctx->g.ny, ctx->W8, ctx->H8, ctx->aP, ctx->x, ctx->rec_snap + (size_t)k * ctx->n,
                                 ctx->rec_scale + k, ctx->rec_t + (size_t)k * img, ctx->st);
        HIPCHK(hipStreamSynchronize(ctx->st));
    }
    // h_new = x + x0 into the alternate buffer (done above when fused_finish),
    // then swap so F.f[SHK_H] is current
    if (!fused_finish) {
        Prof pr(ctx, P_FIN, 3 * nb);
        if (x0 == h_prev) {
            launch::cg_finish((int)ctx->n, ctx->x, h_prev, ctx->h_alt, relax, ctx->st);
        } else {
            launch::cg_finish((int)ctx->n, ctx->x, x0, ctx->h_alt, 0.0, ctx->st);
            if (relax > 0.0) launch::relax(ctx->n, ctx->h_alt, h_prev, relax, ctx->st);
        }
    }
    ctx->F.f[SHK_H] = ctx->h_alt;
    ctx->h_alt = h_prev;
    if (record) {
        launch::nn_target(ctx->g.nx, ctx->g.ny, ctx->W8, ctx->H8, ctx->aP, ctx->F.f[SHK_H], h_prev, ctx->sc,
                          ctx->rec_t + slot, ctx->st);
        ctx->rec_count += 1;
        HIPCHK(hipStreamSynchronize(ctx->st));
    }
    if (ctx->prof) prof_collect(ctx);
    return conclude(ctx, sst, stats);
}

// The field sliding laws read their staggered coefficient.
int check_sliding(shk_ctx* ctx) {
    const int law = ctx->m.sliding_law;
    if ((law == SHK_SLIDING_LINEAR || law == SHK_SLIDING_REGULARIZED_COULOMB_FIELD ||
         law == SHK_SLIDING_REGULARIZED_COULOMB_V0) && !ctx->slide_cx)
        return fail(ctx, SHK_ERR_STATE, "this sliding law needs its coefficient: call shk_set_sliding_coefficient");
    return SHK_OK;
}

// The derived-field refresh (mode 0 after a solve, mode 1 of the initial
// conditions).  full == false writes only N, K and mdot: what the next
// assembly reads when another Picard iteration follows.
void derived_launch(shk_ctx* ctx, int mode, bool full) {
    if (ctx->derived_two_pass)
        launch::derived(ctx->G, ctx->P, ctx->F, nullptr, mode, 0, ctx->partials, ctx->sc, ctx->st);
    else
        launch::derived_march(ctx->P, ctx->F, mode, full || !ctx->lean_refresh, ctx->st);
}

int derived(shk_ctx* ctx, bool full) {
    if (int rc = check_sliding(ctx)) return rc;
    Prof pr(ctx, P_DER, 0);
    derived_launch(ctx, 0, full);
    return SHK_OK;
}

// The Picard test of the current head against h_prev (src/elliptic_solver.jl:
// 153-169), the flag read back into sc_host->picard_converged.
int picard_test(shk_ctx* ctx, const double* h_prev, double tol) {
    ctx->sc_host->picard_tol = tol;
    // only picard_tol changes; write that field alone
    HIPCHK(hipMemcpyAsync(&ctx->sc->picard_tol, &ctx->sc_host->picard_tol, sizeof(double), hipMemcpyHostToDevice,
                          ctx->st));
    {
        Prof pr(ctx, P_DER, 0);
        launch::picard_check(ctx->P, ctx->F.f[SHK_H], h_prev, ctx->partials, ctx->sc, ctx->st);
    }
    return read_scalars(ctx);
}

template <class T>
int dalloc(shk_ctx* ctx, T** p, size_t count) {
    HIPCHK(hipMalloc((void**)p, count * sizeof(T) + 256));
    HIPCHK(hipMemsetAsync(*p, 0, count * sizeof(T) + 256, ctx->st));
    return SHK_OK;
}

}  // namespace

// ================================================ allocation-free calls
//
// Device memory that shk_residual_norms_noalloc uses beyond its inputs and
// outputs: the per-workgroup partials of its two sums of squares, 2 x
// kNoallocMaxBlocks doubles (512 KiB), one copy per GPU for the life of the
// process.  Host code never passes `scratch` itself to a kernel (on the host
// the name stands for no device address): it takes the device address once
// per device with hipGetSymbolAddress and uses that.
constexpr int kNoallocMaxBlocks = 32768;
static __device__ __align__(256) double scratch[2 * kNoallocMaxBlocks];

namespace {
constexpr int kMaxDevices = 64;
double* scratch_dev[kMaxDevices] = {};
std::mutex scratch_mu;
thread_local std::string noctx_error = "null context";

// The device address of `scratch` on the current device, taken on the first
// call on that device (with it current) and kept.
int scratch_on_current_device(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        noctx_error = "no current HIP device (or more than 64 devices)";
        return SHK_ERR_HIP;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            noctx_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SHK_ERR_HIP;
        }
        scratch_dev[dev] = (double*)p;
    }
    *out = scratch_dev[dev];
    return SHK_OK;
}
}  // namespace

// ============================================================= C ABI

extern "C" {

int32_t shk_abi_version(void) { return SHK_ABI_VERSION; }

const char* shk_last_error(const shk_ctx* ctx) { return ctx ? ctx->err.c_str() : noctx_error.c_str(); }

int shk_create(const shk_grid* grid, const shk_params* params, const shk_model* model, const shk_solver_cfg* cfg,
               shk_ctx** out) {
    if (!out || !grid || !params || !model || !cfg) return SHK_ERR_ARG;
    *out = nullptr;
    if (grid->nx < 2 || grid->ny < 2) return SHK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SHK_ERR_NODEVICE;
    shk_ctx* ctx = new shk_ctx();
    ctx->g = *grid;
    ctx->p = *params;
    ctx->m = *model;
    int rc = check_cfg(ctx, cfg);
    if (rc) { delete ctx; return rc; }
    ctx->cfg = *cfg;
    HIPCHK(hipGetDevice(&ctx->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::string arch = prop.gcnArchName;
        delete ctx;
        (void)arch;
        return SHK_ERR_NODEVICE;
    }
    HIPCHK(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
    if (!getenv("SHK_NO_SIDE_INVERSE")) HIPCHK(hipStreamCreateWithFlags(&ctx->st_side, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ctx->inv_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ctx->inv_done, hipEventDisableTiming));
    ctx->use_graphs = getenv("SHK_NO_GRAPHS") == nullptr;
    ctx->use_fused = getenv("SHK_NO_FUSED") == nullptr;
    ctx->use_defer = getenv("SHK_NO_DEFER") == nullptr;
    ctx->use_adaptive = getenv("SHK_FIXED_BATCH") == nullptr;
    if (getenv("SHK_MG_JACOBI")) ctx->mg_rb = 0;
    if (const char* e = getenv("SHK_BATCH_MARGIN")) ctx->batch_margin = atoi(e);
    ctx->derived_two_pass = getenv("SHK_DERIVED_TWO_PASS") != nullptr;
    ctx->lean_refresh = getenv("SHK_NO_LEAN_REFRESH") == nullptr;
    ctx->asm_start = getenv("SHK_NO_ASM_START") == nullptr;
    ctx->pair_x = getenv("SHK_NO_PAIR_X") == nullptr;
    const size_t nx = grid->nx, ny = grid->ny;
    ctx->n = nx * ny;
    ctx->nxf = (nx + 1) * ny;
    ctx->nyf = nx * (ny + 1);
    fill_phys(ctx);
    ctx->G = make_geom((int)nx, (int)ny);
    ctx->W8 = (int)((nx + 63) / 64 * 64);
    ctx->H8 = (int)((ny + 31) / 32 * 32);
    for (int f = 0; f < SHK_NUM_FIELDS; ++f)
        if ((rc = dalloc(ctx, &ctx->F.f[f], field_count(ctx, f)))) return rc;
    if ((rc = dalloc(ctx, &ctx->h_alt, ctx->n))) return rc;
    // State(g): mask GROUNDED (0), valid_x/valid_y = 1
    {
        std::vector<double> ones(std::max(ctx->nxf, ctx->nyf), 1.0);
        HIPCHK(hipMemcpyAsync(ctx->F.f[SHK_VALID_X], ones.data(), ctx->nxf * 8, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipMemcpyAsync(ctx->F.f[SHK_VALID_Y], ones.data(), ctx->nyf * 8, hipMemcpyHostToDevice, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
    }
    double** d64[] = {&ctx->aP, &ctx->aE, &ctx->aN, &ctx->rhs, &ctx->x, &ctx->r, &ctx->z, &ctx->pA,
                      &ctx->pB, &ctx->Ap, &ctx->s, &ctx->cA, &ctx->cB, &ctx->d64, &ctx->r2, &ctx->cP};
    for (auto pp : d64)
        if ((rc = dalloc(ctx, pp, ctx->n))) return rc;
    float** d32[] = {&ctx->aP32, &ctx->aE32, &ctx->aN32, &ctx->b32, &ctx->x32, &ctx->r32, &ctx->z32,
                     &ctx->pA32, &ctx->pB32, &ctx->Ap32, &ctx->s32, &ctx->cA32, &ctx->cB32, &ctx->r2_32,
                     &ctx->dg32};
    for (auto pp : d32)
        if ((rc = dalloc(ctx, pp, ctx->n))) return rc;
    Geom cells = make_geom((int)nx, (int)ny, 1);
    size_t maxblocks = (size_t)std::max(ctx->G.nblocks, cells.nblocks);
    for (int H = 1; H < 8; ++H) maxblocks = std::max(maxblocks, (size_t)fused_geom(ctx->G, H).nblocks);
    const size_t nparts = 2 * maxblocks + 64;
    if ((rc = dalloc(ctx, &ctx->partials, nparts))) return rc;
    if ((rc = dalloc(ctx, &ctx->partials2, nparts))) return rc;
    HIPCHK(hipMalloc((void**)&ctx->sc, sizeof(Scalars)));
    HIPCHK(hipHostMalloc((void**)&ctx->sc_host, sizeof(Scalars), hipHostMallocDefault));
    memset(ctx->sc_host, 0, sizeof(Scalars));
    ctx->sc_host->lmin = 1.0;  // ChebyshevPreconditioner placeholder bounds (src/preconditioner.jl:185)
    ctx->sc_host->lmax = 1.0;
    HIPCHK(hipMemcpy(ctx->sc, ctx->sc_host, sizeof(Scalars), hipMemcpyHostToDevice));
    ctx->prof_tab.resize(P_NUM);
    for (int i = 0; i < P_NUM; ++i) ctx->prof_tab[i].name = kProfNames[i];
    HIPCHK(hipStreamSynchronize(ctx->st));
    *out = ctx;
    return SHK_OK;
}

int shk_destroy(shk_ctx* ctx) {
    if (!ctx) return SHK_ERR_ARG;
    if (ctx->st) hipStreamSynchronize(ctx->st);
    if (ctx->st_side) hipStreamSynchronize(ctx->st_side);
    shkio::destroy(ctx->io);  // joins a running checkpoint writer
    for (int f = 0; f < SHK_NUM_FIELDS; ++f) hipFree(ctx->F.f[f]);
    void* ptrs[] = {ctx->h_alt, ctx->moulin, ctx->slide_cx, ctx->slide_cy, ctx->r2, ctx->r2_32, ctx->cP, ctx->dg32, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->x, ctx->r, ctx->z,
                    ctx->pA, ctx->pB, ctx->Ap, ctx->s, ctx->cA, ctx->cB, ctx->d64, ctx->aP32, ctx->aE32,
                    ctx->aN32, ctx->b32, ctx->x32, ctx->r32, ctx->z32, ctx->pA32, ctx->pB32, ctx->Ap32,
                    ctx->s32, ctx->cA32, ctx->cB32, ctx->partials, ctx->partials2, ctx->sc};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (ctx->sc_host) hipHostFree(ctx->sc_host);
    for (auto& e : ctx->graphs) {
        hipGraphExecDestroy(e.exec);
        hipGraphDestroy(e.graph);
    }
    for (auto& L : ctx->mg) {
        hipFree(L.xa);
        if (L.owned) {
            float* ps[] = {L.b, L.xb};
            for (float* p : ps) hipFree(p);
            if (!L.packed) {
                float* qs[] = {L.c, L.e, L.n, L.D};
                for (float* p : qs) hipFree(p);
            }
        }
    }
    if (ctx->mg_pack) hipFree(ctx->mg_pack);
    for (uint16_t* q : {ctx->mg_c16, ctx->mg_e16, ctx->mg_n16, ctx->mg_D16})
        if (q) hipFree(q);  // holds mg_inv
    if (ctx->net) shknn::release_stream(ctx->net, ctx->st);
    void* nnp[] = {ctx->nn_x, ctx->nn_e, ctx->nn_Ae, ctx->hw, ctx->rec_snap, ctx->rec_scale};
    for (void* p : nnp)
        if (p) hipFree(p);
    for (auto& pe : ctx->pending) { hipEventDestroy(pe.a); hipEventDestroy(pe.b); }
    for (auto e : ctx->ev_pool) hipEventDestroy(e);
    if (ctx->st) hipStreamDestroy(ctx->st);
    if (ctx->st_side) hipStreamDestroy(ctx->st_side);
    if (ctx->inv_fork) hipEventDestroy(ctx->inv_fork);
    if (ctx->inv_done) hipEventDestroy(ctx->inv_done);
    delete ctx;
    return SHK_OK;
}

int shk_set_solver_cfg(shk_ctx* ctx, const shk_solver_cfg* cfg) {
    if (!ctx || !cfg) return SHK_ERR_ARG;
    int rc = check_cfg(ctx, cfg);
    if (rc) return rc;
    ctx->cfg = *cfg;
    return SHK_OK;
}

int shk_set_params(shk_ctx* ctx, const shk_params* params) {
    if (!ctx || !params) return SHK_ERR_ARG;
    ctx->p = *params;
    fill_phys(ctx);
    return SHK_OK;
}

int shk_set_model(shk_ctx* ctx, const shk_model* model) {
    if (!ctx || !model) return SHK_ERR_ARG;
    if (model->kfs < 0 || model->kfs > 1 || model->gap_scheme < 0 || model->gap_scheme > 1 ||
        model->sliding_law < SHK_SLIDING_REGULARIZED_COULOMB || model->sliding_law > SHK_SLIDING_REGULARIZED_COULOMB_V0)
        return fail(ctx, SHK_ERR_ARG, "unknown k-face scheme, gap scheme or sliding law");
    if (model->sliding_law == SHK_SLIDING_REGULARIZED_COULOMB_V0 && !(model->sliding_v0 >= 0.0))
        return fail(ctx, SHK_ERR_ARG, "RegularizedCoulombV0SlidingLaw: v0 must be >= 0");
    ctx->m = *model;
    fill_phys(ctx);
    return SHK_OK;
}

int shk_set_sliding_coefficient(shk_ctx* ctx, const double* C_host, size_t count) {
    if (!ctx || !C_host || count != ctx->n)
        return fail(ctx, SHK_ERR_ARG, "the sliding coefficient must hold nx*ny cell values");
    if (!ctx->slide_cx) {
        HIPCHK(hipMalloc((void**)&ctx->slide_cx, ctx->nxf * 8 + 256));
        HIPCHK(hipMalloc((void**)&ctx->slide_cy, ctx->nyf * 8 + 256));
    }
    HIPCHK(hipMemcpyAsync(ctx->d64, C_host, count * 8, hipMemcpyHostToDevice, ctx->st));
    launch::stagger(ctx->g.nx, ctx->g.ny, ctx->d64, ctx->slide_cx, ctx->slide_cy, ctx->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    fill_phys(ctx);
    return SHK_OK;
}

int shk_upload(shk_ctx* ctx, int32_t field, const double* host, size_t count) {
    if (!ctx || !host) return fail(ctx, SHK_ERR_ARG, "null argument");
    const size_t want = field_count(ctx, field);
    if (!want) return fail(ctx, SHK_ERR_ARG, "unknown field id " + std::to_string(field));
    if (count != want)
        return fail(ctx, SHK_ERR_ARG, "field " + std::to_string(field) + " expects " + std::to_string(want) +
                                          " values, got " + std::to_string(count));
    HIPCHK(hipMemcpyAsync(ctx->F.f[field], host, count * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_download(shk_ctx* ctx, int32_t field, double* host, size_t count) {
    if (!ctx || !host) return fail(ctx, SHK_ERR_ARG, "null argument");
    const size_t want = field_count(ctx, field);
    if (!want) return fail(ctx, SHK_ERR_ARG, "unknown field id " + std::to_string(field));
    if (count != want)
        return fail(ctx, SHK_ERR_ARG, "field " + std::to_string(field) + " expects " + std::to_string(want) +
                                          " values, got " + std::to_string(count));
    HIPCHK(hipMemcpyAsync(host, ctx->F.f[field], count * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

// ---------------------------------------------------------------- snapshots and checkpoints

namespace {

// the reference State's field names, in shk_field order (src/state.jl:13-57)
const char* const kFieldNames[SHK_NUM_FIELDS] = {
    "h",    "pw",   "po",    "b",      "beta",   "abs_ub", "mdot",    "shear",   "potential", "sensible",
    "Re",   "K",    "G",     "zb",     "zs",     "H",      "ieb",     "lambda",  "A_visc",    "N",
    "mask", "dhdx", "q_x",   "Re_x",   "b_x",    "ub_x",   "taub_x",  "dpwdx",   "valid_x",   "dhdy",
    "q_y",  "Re_y", "b_y",   "ub_y",   "taub_y", "dpwdy",  "valid_y"};

void field_descs(const shk_ctx* c, shkio::FieldDesc* d) {
    for (int f = 0; f < SHK_NUM_FIELDS; ++f) {
        d[f].name = kFieldNames[f];
        d[f].rows = c->g.nx + (f > SHK_MASK && f <= SHK_VALID_X ? 1 : 0);
        d[f].cols = c->g.ny + (f > SHK_VALID_X ? 1 : 0);
        d[f].dev = c->F.f[f];
    }
}

int io_ready(shk_ctx* ctx) {
    if (ctx->io) return SHK_OK;
    std::string e;
    ctx->io = shkio::create(e);
    return ctx->io ? SHK_OK : fail(ctx, SHK_ERR_HIP, e);
}

int snapshot_all(shk_ctx* ctx) {
    shkio::FieldDesc d[SHK_NUM_FIELDS];
    field_descs(ctx, d);
    int ids[SHK_NUM_FIELDS];
    for (int f = 0; f < SHK_NUM_FIELDS; ++f) ids[f] = f;
    std::string e;
    const int rc = shkio::begin(ctx->io, ctx->st, SHK_NUM_FIELDS, ids, d, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

}  // namespace

int shk_snapshot(shk_ctx* ctx, const int32_t* field_ids, int32_t nfields) {
    if (!ctx || nfields < 0) return fail(ctx, SHK_ERR_ARG, "bad arguments");
    if (int rc = io_ready(ctx)) return rc;
    if (!field_ids || nfields == 0) return snapshot_all(ctx);
    shkio::FieldDesc d[SHK_NUM_FIELDS];
    field_descs(ctx, d);
    std::vector<int> ids(field_ids, field_ids + nfields);
    std::string e;
    const int rc = shkio::begin(ctx->io, ctx->st, nfields, ids.data(), d, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_snapshot_fetch(shk_ctx* ctx, int32_t field, double* host, size_t count) {
    if (!ctx || !host) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (!ctx->io) return fail(ctx, SHK_ERR_STATE, "no snapshot has been taken");
    std::string e;
    const int rc = shkio::fetch(ctx->io, field, host, count, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_checkpoint_save(shk_ctx* ctx, const char* path, int64_t t, double total_time, int32_t async) {
    if (!ctx || !path || !*path) return fail(ctx, SHK_ERR_ARG, "a checkpoint needs a path");
    if (int rc = io_ready(ctx)) return rc;
    if (int rc = snapshot_all(ctx)) return rc;
    std::string e;
    const int rc = shkio::write_checkpoint(ctx->io, path, ctx->g.nx, ctx->g.ny, ctx->g.lx, ctx->g.ly, t, total_time,
                                           async != 0, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_checkpoint_wait(shk_ctx* ctx) {
    if (!ctx) return SHK_ERR_ARG;
    if (!ctx->io) return SHK_OK;
    std::string e;
    const int rc = shkio::wait(ctx->io, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_checkpoint_load(shk_ctx* ctx, const char* path, int64_t* t, double* total_time) {
    if (!ctx || !path || !t || !total_time) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (int rc = io_ready(ctx)) return rc;
    shkio::FieldDesc d[SHK_NUM_FIELDS];
    field_descs(ctx, d);
    std::string e;
    const int rc = shkio::read_checkpoint(ctx->io, ctx->st, path, ctx->g.nx, ctx->g.ny, SHK_NUM_FIELDS, d, t,
                                          total_time, e);
    if (rc) return fail(ctx, rc, e);
    ctx->have_system = false;  // the assembled system belongs to the State that was replaced
    return SHK_OK;
}

int shk_field_stats(shk_ctx* ctx, int32_t field, double* out) {
    if (!ctx || !out) return fail(ctx, SHK_ERR_ARG, "null argument");
    const size_t n = field_count(ctx, field);
    if (!n) return fail(ctx, SHK_ERR_ARG, "unknown field id " + std::to_string(field));
    if (int rc = io_ready(ctx)) return rc;
    std::string e;
    const int rc = shkio::field_stats(ctx->io, ctx->st, ctx->F.f[field], n, out, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_generate_glacier(shk_ctx* ctx, const shk_glacier_spec* spec) {
    if (!ctx || !spec) return fail(ctx, SHK_ERR_ARG, "null argument");
    const shk_glacier_spec& s = *spec;
    const int64_t interior = (int64_t)(ctx->g.nx - 2) * (ctx->g.ny - 2);
    if (ctx->g.nx < 2 || ctx->g.ny < 2) return fail(ctx, SHK_ERR_ARG, "the generator needs nx, ny >= 2");
    if (!(s.bed_amp >= 0.0) || !(s.corr_len >= 0.0) || !(s.trough_width >= 0.0) || !std::isfinite(s.trough_depth) ||
        !std::isfinite(s.bed_amp) || !std::isfinite(s.corr_len) || !std::isfinite(s.trough_width))
        return fail(ctx, SHK_ERR_ARG, "bed_amp, corr_len and trough_width must be finite and >= 0");
    if (s.corr_len / (sqrt(2.0) * std::min(ctx->g.dx, ctx->g.dy)) > 1e6)
        return fail(ctx, SHK_ERR_ARG, "corr_len spans more than a million cells");
    if (s.n_moulins < 0 || (s.n_moulins > 0 && (ctx->g.nx < 3 || ctx->g.ny < 3 || s.n_moulins > interior)))
        return fail(ctx, SHK_ERR_ARG, "n_moulins must be between 0 and the number of interior cells");
    if (s.n_moulins > 0 && !(s.q_max >= s.q_min))
        return fail(ctx, SHK_ERR_ARG, "q_max must be >= q_min");
    if (s.n_moulins > 0 && !ctx->moulin) HIPCHK(hipMalloc((void**)&ctx->moulin, ctx->n * 8));
    std::string e;
    const int rc = shkgen::generate(ctx->g, s, ctx->F, ctx->d64, ctx->r, s.n_moulins > 0 ? ctx->moulin : nullptr,
                                    ctx->st, e);
    return rc ? fail(ctx, rc, e) : SHK_OK;
}

int shk_set_initial_conditions(shk_ctx* ctx) {
    if (!ctx) return SHK_ERR_ARG;
    if (int rc = check_sliding(ctx)) return rc;
    {
        Prof pr(ctx, P_IC, 0);
        launch::ic_cells(ctx->P, ctx->F, ctx->st);
        launch::stagger_b(ctx->P, ctx->F, ctx->st);
        derived_launch(ctx, 1, true);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (ctx->prof) prof_collect(ctx);
    return SHK_OK;
}

int shk_solve_elliptic(shk_ctx* ctx, shk_solve_stats* stats) {
    if (!ctx) return SHK_ERR_ARG;
    int rc = solve(ctx, 0, 0.0, -1.0, stats);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_solve_parabolic(shk_ctx* ctx, double dt, shk_solve_stats* stats) {
    if (!ctx || !(dt > 0)) return fail(ctx, SHK_ERR_ARG, "dt must be > 0");
    int rc = solve(ctx, 1, dt, -1.0, stats);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_update_derived(shk_ctx* ctx) {
    if (!ctx) return SHK_ERR_ARG;
    int rc = derived(ctx, true);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_picard_iteration(shk_ctx* ctx, double relax_alpha, shk_solve_stats* stats) {
    if (!ctx) return SHK_ERR_ARG;
    int rc = solve(ctx, 0, 0.0, relax_alpha, stats);
    if (rc) return rc;
    rc = derived(ctx, true);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_picard_loop(shk_ctx* ctx, int32_t iters, double tol, int32_t check_every, double relax_alpha,
                    int32_t* converged, int32_t* last_iter, int64_t* cg_iters) {
    if (!ctx || iters < 1 || check_every < 1) return fail(ctx, SHK_ERR_ARG, "iters and check_every must be >= 1");
    int32_t conv = 0, last = 0;
    struct FloorReset {
        shk_ctx* c;
        ~FloorReset() {
            c->atol_floor = 0.0;
            c->picard_it = 0;
            c->picard_fuse = false;
        }
    } floor_reset{ctx};
    ctx->atol_floor = 0.0;
    // the tolerance of the Picard test, on the device for the whole loop
    ctx->sc_host->picard_tol = tol;
    HIPCHK(hipMemcpyAsync(&ctx->sc->picard_tol, &ctx->sc_host->picard_tol, sizeof(double), hipMemcpyHostToDevice,
                          ctx->st));
    for (int32_t it = 1; it <= iters; ++it) {
        shk_solve_stats st;
        const int check = (it % check_every == 0) || (it == iters);
        ctx->picard_it = it;
        ctx->picard_fuse = check;
        ctx->picard_fused = false;
        int rc = solve(ctx, 0, 0.0, relax_alpha, &st);
        ctx->picard_it = 0;
        ctx->picard_fuse = false;
        if (rc) return rc;
        if (it == 1 && (ctx->cfg.flags & SHK_CFG_PICARD_ATOL) && std::isfinite(st.rnorm0))
            ctx->atol_floor = ctx->cfg.rtol * st.rnorm0;
        if (cg_iters) *cg_iters += st.niter;
        // The test reads only the heads, so it runs before the refresh: the
        // refresh after the loop's last iteration writes every field, the
        // others only what the next assembly reads.  The solve takes it in
        // its epilogue where it can (the flag then came with its last read).
        bool done = false;
        if (check) {
            if (!ctx->picard_fused) {
                rc = picard_test(ctx, ctx->h_alt /* previous h */, tol);
                if (rc) return rc;
            }
            // a solve that stopped short (itmax, zero curvature) leaves a head
            // change that says nothing about Picard convergence
            done = ctx->sc_host->picard_converged && !st.zero_curvature && !st.tired;
        }
        rc = derived(ctx, done || it == iters);
        if (rc) return rc;
        if (done) {
            conv = 1;
            last = it;
            break;
        }
    }
    if (!conv) last = iters;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (converged) *converged = conv;
    if (last_iter) *last_iter = last;
    return SHK_OK;
}

int shk_solve_log(shk_ctx* ctx, shk_solve_stats* out, int32_t max_entries, int32_t* n_entries) {
    if (!ctx || !n_entries || max_entries < 0 || (max_entries > 0 && !out))
        return fail(ctx, SHK_ERR_ARG, "shk_solve_log: bad argument");
    const size_t n = ctx->solve_log.size();
    const size_t k = std::min(n, (size_t)max_entries);
    for (size_t i = 0; i < k; ++i) out[i] = ctx->solve_log[n - k + i];
    *n_entries = (int32_t)k;
    ctx->solve_log.clear();
    return SHK_OK;
}

int shk_parabolic_step(shk_ctx* ctx, double dt, int64_t* cg_iters) {
    if (!ctx || !(dt > 0)) return fail(ctx, SHK_ERR_ARG, "dt must be > 0");
    shk_solve_stats st;
    int rc = solve(ctx, 1, dt, -1.0, &st);
    if (rc) return rc;
    if (cg_iters) *cg_iters += st.niter;
    rc = derived(ctx, true);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_step_b(shk_ctx* ctx, double dt) {
    if (!ctx) return SHK_ERR_ARG;
    {
        Prof pr(ctx, P_GAP, 0);
        launch::gap(ctx->P, ctx->F, dt, ctx->st);
    }
    {
        Prof pr(ctx, P_STAG, 0);
        launch::stagger_b(ctx->P, ctx->F, ctx->st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (ctx->prof) prof_collect(ctx);
    return SHK_OK;
}

int shk_set_moulin_field(shk_ctx* ctx, const double* host, size_t count) {
    if (!ctx || !host || count != ctx->n) return fail(ctx, SHK_ERR_ARG, "moulin field must hold nx*ny values");
    if (!ctx->moulin) HIPCHK(hipMalloc((void**)&ctx->moulin, ctx->n * 8));
    HIPCHK(hipMemcpyAsync(ctx->moulin, host, count * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_update_ieb(shk_ctx* ctx, int32_t mode, double uniform) {
    if (!ctx || mode < 0 || mode > 2) return fail(ctx, SHK_ERR_ARG, "ieb mode must be 0, 1 or 2");
    if (mode == 0) return SHK_OK;
    if (mode == 2 && !ctx->moulin) return fail(ctx, SHK_ERR_STATE, "ieb mode 2 needs shk_set_moulin_field first");
    {
        Prof pr(ctx, P_IEB, 0);
        launch::fill_ieb((int)ctx->n, ctx->F.f[SHK_IEB], uniform, mode == 2 ? ctx->moulin : nullptr, ctx->st);
    }
    HIPCHK(hipGetLastError());
    return SHK_OK;
}

int shk_step(shk_ctx* ctx, double dt, int32_t ieb_mode, double ieb_uniform, int32_t iters, double tol,
             int32_t check_every, double relax_alpha, int32_t* converged, int32_t* last_iter, int64_t* cg_iters) {
    if (!ctx) return SHK_ERR_ARG;
    int rc = shk_update_ieb(ctx, ieb_mode, ieb_uniform);
    if (rc) return rc;
    if (ctx->p.e_v == 0.0) {
        rc = shk_picard_loop(ctx, iters, tol, check_every, relax_alpha, converged, last_iter, cg_iters);
    } else {
        rc = shk_parabolic_step(ctx, dt, cg_iters);
        if (converged) *converged = 1;
        if (last_iter) *last_iter = 0;
    }
    if (rc) return rc;
    return shk_step_b(ctx, dt);
}

int shk_assemble(shk_ctx* ctx, int32_t parabolic, double dt) {
    if (!ctx) return SHK_ERR_ARG;
    if (parabolic && !(dt > 0)) return fail(ctx, SHK_ERR_ARG, "dt must be > 0");
    launch::assemble<double>(ctx->G, ctx->P, ctx->F, parabolic, dt, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->st,
                             ctx->cP);
    ctx->have_system = true;
    ctx->system_parabolic = parabolic;
    ctx->system_dt = dt;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_apply_operator(shk_ctx* ctx, const double* x_host, double* y_host) {
    if (!ctx || !x_host || !y_host) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    HIPCHK(hipMemcpyAsync(ctx->d64, x_host, ctx->n * 8, hipMemcpyHostToDevice, ctx->st));
    {
        Prof pr(ctx, P_MATVEC, 5.0 * ctx->n * 8);
        launch::matvec<double>(ctx->G, ctx->aP, ctx->aE, ctx->aN, ctx->d64, ctx->Ap, ctx->st);
    }
    HIPCHK(hipMemcpyAsync(y_host, ctx->Ap, ctx->n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_update_chebyshev_bounds(shk_ctx* ctx, const double* rhs_host, double* lambda_min, double* lambda_max,
                                int32_t* warned) {
    if (!ctx || !rhs_host) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    if (ctx->cfg.precond != SHK_PRECOND_CHEBYSHEV)
        return fail(ctx, SHK_ERR_STATE, "shk_update_chebyshev_bounds: the configured preconditioner is not Chebyshev");
    HIPCHK(hipMemcpyAsync(ctx->d64, rhs_host, ctx->n * 8, hipMemcpyHostToDevice, ctx->st));
    // the solver's own vectors as the estimate's scratch, the given rhs as its start
    Work<double> w{ctx->aP, ctx->aE, ctx->aN, ctx->aP, ctx->d64, ctx->x, ctx->r, ctx->z, ctx->pA, ctx->pB,
                   ctx->Ap, ctx->s, ctx->cA, ctx->cB, ctx->r2};
    cheb_refresh<double>(ctx, w, ctx->cfg.cheb_nsteps, ctx->cfg.cheb_degree);
    HIPCHK(hipGetLastError());
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (lambda_min) *lambda_min = ctx->sc_host->lmin;
    if (lambda_max) *lambda_max = ctx->sc_host->lmax;
    if (warned) *warned = ctx->sc_host->bounds_warning;
    return SHK_OK;
}

int shk_residual_norms(shk_ctx* ctx, const double* x_host, double* norms) {
    if (!ctx || !x_host || !norms) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    HIPCHK(hipMemcpyAsync(ctx->d64, x_host, ctx->n * 8, hipMemcpyHostToDevice, ctx->st));
    launch::residual_norms(ctx->G, ctx->aP, ctx->aE, ctx->aN, ctx->rhs, ctx->d64, nullptr, ctx->partials,
                           &ctx->sc->tres, ctx->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(norms, &ctx->sc->tres, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_residual_norms_noalloc(int32_t nx, int32_t ny, const double* aP, const double* aE, const double* aN,
                               const double* b, const double* x, double* out_dev, void* stream) {
    if (nx < 2 || ny < 2) {
        noctx_error = "shk_residual_norms_noalloc: nx and ny must be >= 2";
        return SHK_ERR_ARG;
    }
    if (!aP || !aE || !aN || !b || !x || !out_dev) {
        noctx_error = "shk_residual_norms_noalloc: null device pointer";
        return SHK_ERR_ARG;
    }
    // the geometry shk_residual_norms runs for this grid (a context's)
    const Geom G = make_geom(nx, ny);
    const int nblk = launch::residual_norms_nblocks(G);  // the true-residual kernel's workgroups
    if (nblk > kNoallocMaxBlocks) {
        noctx_error = "shk_residual_norms_noalloc: a " + std::to_string(nx) + " x " + std::to_string(ny) +
                      " grid needs " + std::to_string(nblk) + " workgroups, more than the " +
                      std::to_string(kNoallocMaxBlocks) + " partials of the static scratch";
        return SHK_ERR_ARG;
    }
    double* part = nullptr;
    const int rc = scratch_on_current_device(&part);
    if (rc) return rc;
    launch::residual_norms(G, aP, aE, aN, b, x, nullptr, part, out_dev, (hipStream_t)stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        noctx_error = std::string("shk_residual_norms_noalloc: ") + hipGetErrorString(e);
        return SHK_ERR_HIP;
    }
    return SHK_OK;
}

int shk_apply_precond(shk_ctx* ctx, const double* r_host, double* z_host) {
    if (!ctx || !r_host || !z_host) return fail(ctx, SHK_ERR_ARG, "null argument");
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    if (ctx->cfg.precond != SHK_PRECOND_MULTIGRID)
        return fail(ctx, SHK_ERR_STATE, "shk_apply_precond: the configured preconditioner is not multigrid");
    // the FP32 flux-form hierarchy of the last assembly, as the solve builds it
    launch::to_f32(ctx->n, ctx->cP, ctx->aP32, ctx->st);
    launch::to_f32(ctx->n, ctx->aE, ctx->aE32, ctx->st);
    launch::to_f32(ctx->n, ctx->aN, ctx->aN32, ctx->st);
    launch::to_f32(ctx->n, ctx->aP, ctx->dg32, ctx->st);
    Work<float> wi{ctx->aP32, ctx->aE32, ctx->aN32, ctx->dg32, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = mg_setup(ctx, wi);
    if (rc) return rc;
    rc = reset_scalars(ctx, 0.0, 0.0, 1);  // the V-cycle kernels run while the done flag is clear
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->d64, r_host, ctx->n * 8, hipMemcpyHostToDevice, ctx->st));
    mg_apply<double>(ctx, ctx->d64, ctx->Ap, -1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z_host, ctx->Ap, ctx->n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return SHK_OK;
}

int shk_download_system(shk_ctx* ctx, double* aP, double* aE, double* aW, double* aN, double* aS, double* rhs) {
    if (!ctx) return SHK_ERR_ARG;
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    const size_t n = ctx->n, nx = ctx->g.nx, ny = ctx->g.ny;
    std::vector<double> e(n), no(n);
    HIPCHK(hipMemcpyAsync(e.data(), ctx->aE, n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipMemcpyAsync(no.data(), ctx->aN, n * 8, hipMemcpyDeviceToHost, ctx->st));
    if (aP) HIPCHK(hipMemcpyAsync(aP, ctx->aP, n * 8, hipMemcpyDeviceToHost, ctx->st));
    if (rhs) HIPCHK(hipMemcpyAsync(rhs, ctx->rhs, n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    // Expand the symmetric face layout back into the reference's five arrays.
    for (size_t iy = 0; iy < ny; ++iy)
        for (size_t ix = 0; ix < nx; ++ix) {
            const size_t c = ix + iy * nx;
            if (aE) aE[c] = e[c];
            if (aN) aN[c] = no[c];
            if (aW) aW[c] = ix > 0 ? e[c - 1] : 0.0;
            if (aS) aS[c] = iy > 0 ? no[c - nx] : 0.0;
        }
    return SHK_OK;
}

int shk_time_kernel(shk_ctx* ctx, int32_t kernel, int32_t reps, double* mean_ms, double* bytes_per_launch) {
    if (!ctx || reps < 1 || !mean_ms || !bytes_per_launch) return fail(ctx, SHK_ERR_ARG, "bad argument");
    if (!ctx->have_system) return fail(ctx, SHK_ERR_STATE, "no assembled system");
    const Geom& G = ctx->G;
    const double n = (double)ctx->n;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    double bytes = 0;
    // the variant the V-cycle-preconditioned solve runs: z in FP32 where its
    // finest level is BF16 and the matvec the two-column kernel
    const bool z32 = ctx->use_defer && ctx->mg_bf16 && cg_wide_ok(G);
    // ... taking its x steps in pairs: the odd and even iterations' kernels in turn
    const bool pair = z32 && ctx->pair_x;
    int calls = 0;
    auto one = [&]() {
        switch (kernel) {
            case 0:
            case 3: launch::matvec<double>(G, ctx->aP, ctx->aE, ctx->aN, ctx->h_alt, ctx->Ap, ctx->st); break;
            case 1: launch::matvec<float>(G, ctx->aP32, ctx->aE32, ctx->aN32, ctx->x32, ctx->Ap32, ctx->st); break;
            case 2:
            case 4: launch::copy(ctx->h_alt, ctx->d64, ctx->n, ctx->st); break;
            case 5:
                if (z32)
                    launch::cg_matvec_deferred(G, ctx->aP, ctx->aE, ctx->aN, (const float*)ctx->z32, ctx->pA,
                                               ctx->pB, ctx->Ap, ctx->partials2, ctx->sc, ctx->partials, 0, ctx->x,
                                               ctx->st, pair ? ((calls++ & 1) ? 2 : 0) : 1);
                else
                    launch::cg_matvec_deferred(G, ctx->aP, ctx->aE, ctx->aN, ctx->z, ctx->pA, ctx->pB, ctx->Ap,
                                               ctx->partials2, ctx->sc, ctx->partials, 0, ctx->x, ctx->st);
                break;
            default: break;
        }
    };
    if (kernel == 0 || kernel == 3) bytes = 5 * n * 8;
    else if (kernel == 1) bytes = 5 * n * 4;
    else if (kernel == 2 || kernel == 4) bytes = 2 * n * 8;
    else if (kernel == 5) bytes = pair ? 8.0 * n * 8 : z32 ? 8.5 * n * 8 : 9 * n * 8;  // pair: the mean of 6.5 and 9.5
    else {
        hipEventDestroy(a);
        hipEventDestroy(b);
        return fail(ctx, SHK_ERR_ARG, "unknown kernel id");
    }
    Scalars saved{};
    if (kernel == 5) {
        // the CG matvec runs only while the solve is live: clear done and the
        // pending flag for the timing, restore the scalars afterwards
        HIPCHK(hipStreamSynchronize(ctx->st));
        HIPCHK(hipMemcpy(&saved, ctx->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
        Scalars live = saved;
        live.done = 0;
        live.gpend = 0;
        live.done_defer = 0;
        live.alpha = 0.0;  // x += 0 p: the traffic of the solve's matvec, x left as it is
        live.alpha_pend = 0.0;
        HIPCHK(hipMemcpy(ctx->sc, &live, sizeof(Scalars), hipMemcpyHostToDevice));
    }
    one();  // warm
    float ms = 0;
    if (kernel >= 3) {
        // cold (3, 4, 5): stream a 1 GiB read before every launch, so L2 and
        // the 256 MB Infinity Cache hold only clean lines of unrelated data;
        // each launch is timed alone
        const size_t nf = (size_t)1 << 27;
        double* flush = nullptr;
        double* sink = nullptr;
        HIPCHK(hipMalloc((void**)&flush, nf * 8));
        HIPCHK(hipMalloc((void**)&sink, 4096 * 8));
        HIPCHK(hipMemsetAsync(flush, 0, nf * 8, ctx->st));
        for (int i = 0; i < reps; ++i) {
            launch::read_flush(nf, flush, sink, ctx->st);
            HIPCHK(hipEventRecord(a, ctx->st));
            one();
            HIPCHK(hipEventRecord(b, ctx->st));
            HIPCHK(hipEventSynchronize(b));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, a, b));
            ms += t;
        }
        hipFree(flush);
        hipFree(sink);
    } else {
        HIPCHK(hipEventRecord(a, ctx->st));
        for (int i = 0; i < reps; ++i) one();
        HIPCHK(hipEventRecord(b, ctx->st));
        HIPCHK(hipEventSynchronize(b));
        HIPCHK(hipEventElapsedTime(&ms, a, b));
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (kernel == 5) HIPCHK(hipMemcpy(ctx->sc, &saved, sizeof(Scalars), hipMemcpyHostToDevice));
    *mean_ms = ms / reps;
    *bytes_per_launch = bytes;
    return SHK_OK;
}

int shk_profile_enable(shk_ctx* ctx, int32_t mask) {
    if (!ctx) return SHK_ERR_ARG;
    if (!mask) prof_collect(ctx);
    ctx->prof = mask != 0;
    ctx->prof_mask = mask == 1 ? 0xffffffffu : (uint32_t)mask;
    return SHK_OK;
}

int shk_profile_read(shk_ctx* ctx, int32_t max_entries, const char** names, int64_t* counts, double* total_ms,
                     double* bytes, int32_t* n_entries) {
    if (!ctx || !n_entries) return SHK_ERR_ARG;
    prof_collect(ctx);
    int k = 0;
    for (int i = 0; i < P_NUM && k < max_entries; ++i, ++k) {
        if (names) names[k] = ctx->prof_tab[i].name;
        if (counts) counts[k] = ctx->prof_tab[i].count;
        if (total_ms) total_ms[k] = ctx->prof_tab[i].ms;
        if (bytes) bytes[k] = ctx->prof_tab[i].bytes;
    }
    *n_entries = k;
    return SHK_OK;
}

int shk_profile_reset(shk_ctx* ctx) {
    if (!ctx) return SHK_ERR_ARG;
    prof_collect(ctx);
    for (auto& e : ctx->prof_tab) { e.count = 0; e.ms = 0; e.bytes = 0; }
    return SHK_OK;
}

// ---- learned warm start (shakti_learned.h) ----

static int nn_buffers(shk_ctx* ctx) {
    if (!ctx->nn_e) {
        HIPCHK(hipMalloc((void**)&ctx->nn_e, ctx->n * 8 + 256));
        HIPCHK(hipMalloc((void**)&ctx->nn_Ae, ctx->n * 8 + 256));
        HIPCHK(hipMalloc((void**)&ctx->hw, ctx->n * 8 + 256));
        HIPCHK(hipMalloc(&ctx->nn_x, (size_t)ctx->W8 * ctx->H8 * 16 + 256));
    }
    return SHK_OK;
}

int shk_attach_unet(shk_ctx* ctx, shk_unet* net) {
    if (!ctx) return SHK_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (ctx->net) shknn::release_stream(ctx->net, ctx->st);
    ctx->net = nullptr;
    // captured learned batches hold the previous network's buffers
    for (size_t i = 0; i < ctx->graphs.size();) {
        auto& e = ctx->graphs[i];
        if (e.net) {
            hipGraphExecDestroy(e.exec);
            hipGraphDestroy(e.graph);
            ctx->graphs.erase(ctx->graphs.begin() + i);
        } else {
            ++i;
        }
    }
    if (!net) return SHK_OK;
    if (!shknn::fits(net, ctx->H8, ctx->W8))
        return fail(ctx, SHK_ERR_ARG,
                    "network capacity too small for one image of this grid (padded to " + std::to_string(ctx->H8) +
                        " x " + std::to_string(ctx->W8) + ")");
    int rc = nn_buffers(ctx);
    if (rc) return rc;
    ctx->net = net;
    shknn::set_stream(net, ctx->st);
    return SHK_OK;
}

int shk_set_learned_precond(shk_ctx* ctx, int32_t every, double stall_ratio) {
    if (!ctx) return SHK_ERR_ARG;
    if (every < 0 || every == 1 || every > 4096)
        return fail(ctx, SHK_ERR_ARG, "learned preconditioner period must be 0 (off) or 2..4096 iterations");
    if (!(stall_ratio >= 0.0) || !(stall_ratio <= 1e3))
        return fail(ctx, SHK_ERR_ARG, "stall ratio must lie in [0, 1000]");
    ctx->nn_every = every;
    ctx->nn_stall = stall_ratio;
    return SHK_OK;
}

static void rec_release(shk_ctx* ctx) {
    if (ctx->rec_snap) hipFree(ctx->rec_snap);
    if (ctx->rec_scale) hipFree(ctx->rec_scale);
    ctx->rec_snap = ctx->rec_scale = nullptr;
    ctx->rec_x = nullptr;
    ctx->rec_t = nullptr;
    ctx->rec_cap = ctx->rec_count = 0;
    ctx->rec_kind = 0;
}

static int record_start(shk_ctx* ctx, void* x_dev, float* t_dev, int32_t capacity, int kind) {
    if (!ctx) return SHK_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    rec_release(ctx);
    if (!x_dev) return SHK_OK;
    if (!t_dev || capacity < 1) return fail(ctx, SHK_ERR_ARG, "record buffers need a target buffer and capacity >= 1");
    int rc = nn_buffers(ctx);
    if (rc) return rc;
    if (kind == 1) {
        HIPCHK(hipMalloc((void**)&ctx->rec_snap, (size_t)capacity * ctx->n * 8 + 256));
        HIPCHK(hipMalloc((void**)&ctx->rec_scale, (size_t)capacity * 8 + 256));
    }
    ctx->rec_x = x_dev;
    ctx->rec_t = t_dev;
    ctx->rec_cap = capacity;
    ctx->rec_count = 0;
    ctx->rec_kind = kind;
    return SHK_OK;
}

int shk_record_pairs(shk_ctx* ctx, void* x_dev, float* t_dev, int32_t capacity) {
    return record_start(ctx, x_dev, t_dev, capacity, 0);
}

int shk_record_cg_pairs(shk_ctx* ctx, void* x_dev, float* t_dev, int32_t capacity) {
    return record_start(ctx, x_dev, t_dev, capacity, 1);
}

int shk_recorded_pairs(shk_ctx* ctx, int32_t* recorded) {
    if (!ctx || !recorded) return SHK_ERR_ARG;
    *recorded = ctx->rec_count;
    return SHK_OK;
}

}  // extern "C"
