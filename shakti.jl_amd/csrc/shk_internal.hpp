// Internal declarations shared by the kernel and API translation units of
// libshakti_hip.so.  Not part of the public ABI (include/shakti_hip.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "shakti_hip.h"

namespace shk {

constexpr double GROUNDED = 0.0, OCEAN = 1.0, LAND = 2.0, OTHER_BASIN = 3.0;  // src/mask.jl:21-41
constexpr double EPS64 = 2.220446049250313e-16;                                // eps(Float64)

constexpr int WAVE = 64;         // CDNA wavefront width
constexpr int WAVES = 4;         // waves per workgroup, stacked along y
constexpr int MAX_CHEB = 32;     // highest supported Chebyshev degree
constexpr int MAX_LANCZOS = 64;  // highest supported bound-estimate step count

// A canonicalised exponent (src/model_parameters.jl:122-134): integer-valued
// exponents take the integer power path.
struct Expo {
    int is_int;
    int k;
    double e;
    int third;  // e is the double nearest 1/3 (the sliding laws' 1/n for n = 3)
};

// Device copy of the State: one pointer per field, in shk_field order.
struct Fields {
    double* f[SHK_NUM_FIELDS];
};

// Scalar state of the solver, resident in device memory.  Written only by the
// single finishing workgroup of a reduction kernel; read by every later kernel.
struct Scalars {
    // Krylov CG recurrence (see k_cg_* kernels)
    double gamma, pNorm2, alpha, beta, rNorm, rNorm0, eps_stop, pAp, atol, rtol;
    int32_t iter, itmax, done, solved, zero_curv, tired;
    // Chebyshev preconditioner: persistent bounds and this solve's coefficients
    double lmin, lmax;
    double cheb_alpha[MAX_CHEB], cheb_beta[MAX_CHEB];
    int32_t bounds_warning, cheb_degree;
    // Lanczos-from-CG bound estimate
    double lz_gamma;
    double lz_alpha[MAX_LANCZOS], lz_beta[MAX_LANCZOS];
    double lz_lmin, lz_lmax;
    // Picard convergence check
    double dmax, hmax, picard_tol;
    int32_t picard_converged, picard_nan;
    // mixed-precision outer loop (FP64 residual in the Jacobi-weighted norm)
    double ir_rnorm, ir_rnorm0, ir_atol, ir_rtol;
    int32_t ir_done, ir_pad;
    // learned warm start: image scale, energy-minimising step, and the
    // preconditioned residual norm at h that the stopping test refers to
    // (0: use the CG start's own norm, as Krylov does)
    double nn_scale, nn_gamma, rnorm0_ref;
    int32_t nn_fallback, nn_pad;
    // learned preconditioner inside flexible CG: image scale and step
    double nn_scale_pc, nn_gamma_pc;
    // r'z partials of the last preconditioner not yet finished (deferred
    // finishing: the next CG matvec sums them)
    int32_t gpend;
    // deferred finishing: the CG matvec's stop decision (converged, tired or
    // breakdown).  The matvec never writes `done`, which its own workgroups
    // test on entry; the kernel after it publishes done from this flag.
    int32_t done_defer;
    // breakdown of the Krylov recurrence: 1 when r'z < 0 or not finite (the
    // preconditioner or the operator is not positive definite; Krylov's
    // sqrt(r'z) throws there), 2 when p'Ap is not finite
    int32_t breakdown, bpad;
    double breakdown_value;  // the offending r'z or p'Ap
    // Gershgorin bound of D^-1 A: 1 + max_i sum_j |a_ij| / a_ii
    double gersh;
    // FP64 residual of the solution a solve returns: ||b - A h||_2, ||b||_2
    double tres, tres_b;
    // unused (reductions finish in a separate kernel); reset with the rest
    uint32_t counter;
    uint32_t pad;
    // CG matvecs taking the x steps in pairs: the step an odd iteration left
    // owed (alpha_pend p, p = xpend_p) while xpend_n is 1
    double alpha_pend;
    const double* xpend_p;
    int32_t xpend_n, xpend_pad;
};

// One level of the multigrid hierarchy as the device sees it, and the small
// levels handled by the single-workgroup tail kernel (level 0 of T is the
// first level below MG_TAIL_CELLS cells, the last one is the coarsest).
struct MGLevelDev {
    int nx, ny;
    const float *c, *e, *n, *D;
    float *b, *xa, *xb;
};
constexpr int MG_MAX_TAIL = 12;
// levels up to this size use the single-phase kernels, larger ones the
// marching kernels (measured faster down to 16k-cell levels at 4096 x 1024)
constexpr long MG_FLAT_CELLS = 4096;
// levels above this many cells take four rows per wave in the two-sweep red-black legs
constexpr long MG_RBN4_CELLS = 1L << 19;
constexpr long MG_TAIL_CELLS = 4096;  // at most 4 cells per thread; levels of ~5.5k cells take ~150 KB of LDS
struct MGTail {
    int nlev;
    float omega, scale;
    const float* inv;
    // c, e, n, D of every tail level and the coarsest inverse, packed in the
    // order the kernel keeps them in LDS: [c | e | n | D] with a stride of
    // `stride` floats (the tail's cell count rounded up to 4), then the inverse
    const float* pack;
    int stride;
    int rb;  // 1: red-black Gauss-Seidel smoothing, 0: damped Jacobi
    // bit l: tail level l reaches level l+1 by the linear transfers (red-black
    // smoothing only), with coarse-correction scale scale_lin; their
    // restriction stages the x-restricted residual rows in LDS after the
    // packed block (hcells floats, the largest such level's coarse width x
    // fine height)
    int lin_mask;
    float scale_lin;
    int hcells;
    MGLevelDev L[MG_MAX_TAIL];
};

// Launch geometry of the row-marching kernels: workgroups of WAVE x WAVES
// threads; each wave walks R consecutive rows of one 64-wide column strip.
struct Geom {
    int nx, ny, R;
    dim3 grid, block;
    int nblocks;
};

Geom make_geom(int nx, int ny, int R_override = 0);
bool wide_ok(const Geom& G);
Geom wide_geom(const Geom& G);
bool cg_wide_ok(const Geom& G);
bool stream_nt(const Geom& G);
Geom fused_geom(const Geom& G, int H);

// Physical constants and scheme switches handed to every physics kernel.
struct Phys {
    int nx, ny;
    double dx, dy, dx2, dy2;
    double rho_w, rho_sw, rho_i, g, nu, n, omega, L, br, lr, ct, cw, p_atm, b_min, b_max, e_v;
    Expo n_exp, n1_exp, inv_n_exp;
    int with_sensible, with_open_by_sliding, kfs, gap_scheme, sliding_law;
    double C, v0;
    // staggered coefficient of the field sliding laws ((nx+1) x ny, nx x (ny+1))
    const double *Cx, *Cy;
};

// Kernel launchers (shk_kernels.hip).  All are asynchronous on `st`.
namespace launch {
void ic_cells(const Phys& P, Fields F, hipStream_t st);
void stagger_b(const Phys& P, Fields F, hipStream_t st);
// cell-centred c onto the faces: outer faces the cell value, inner the mean
void stagger(int nx, int ny, const double* c, double* fx, double* fy, hipStream_t st);
// mode 0: refresh after a solve (pw from h); mode 1: initial conditions (pw as stored).
// check != 0 also reduces max|h - h_prev| and max|h| and sets the Picard flag.
void derived(const Geom& G, const Phys& P, Fields F, const double* h_prev, int mode, int check, double* partials,
             Scalars* sc, hipStream_t st);
// The same refresh in one marching pass (mode as above, no check): full writes
// every field; otherwise only N, K and mdot, for a Picard iteration that
// another one follows.
void derived_march(const Phys& P, Fields F, int mode, bool full, hipStream_t st);
// The Picard test alone: max|h - h_prev| and max|h|, then the Picard flag.
void picard_check(const Phys& P, const double* h, const double* h_prev, double* partials, Scalars* sc,
                  hipStream_t st);
void gap(const Phys& P, Fields F, double dt, hipStream_t st);
void fill_ieb(int n, double* ieb, double uniform, const double* moulin, hipStream_t st);
void zero(double* p, size_t n, hipStream_t st);
void copy(const double* a, double* b, size_t n, hipStream_t st);

// Low-precision copies the assembly can write beside the FP64 system: the
// FP32 flux form (c = cP, e = aE, n = aN, d = aP, each rounded to nearest) and
// those FP32 values rounded to BF16 (nearest even); a null pointer skips a set.
struct AsmLow {
    float *c = nullptr, *e = nullptr, *n = nullptr, *d = nullptr;
    uint16_t *c16 = nullptr, *e16 = nullptr, *n16 = nullptr, *d16 = nullptr;
};
// The start of a CG solve from x0 = h, formed by the FP64 assembly itself
// (k_cg_init without Jacobi): r = rhs - A h, x = 0, p = 0.
struct AsmInit {
    double *r = nullptr, *x = nullptr, *p = nullptr;
};
template <class T>
void assemble(const Geom& G, const Phys& P, Fields F, int parabolic, double dt, T* aP, T* aE, T* aN, T* rhs,
              hipStream_t st, double* cP = nullptr, const AsmLow& lo = AsmLow{}, const AsmInit& in = AsmInit{});
template <class T>
void matvec(const Geom& G, const T* aP, const T* aE, const T* aN, const T* x, T* y, hipStream_t st);

// Lanczos-from-CG bound estimate and the Chebyshev coefficients; sym != 0
// runs it in the D-inner product (D = dg), where D^-1 A is self-adjoint.
template <class T>
void lanczos_init(const Geom& G, const T* rhs, const T* dg, int sym, T* r, T* p0, double* partials, Scalars* sc,
                  hipStream_t st);
template <class T>
void lanczos_step_a(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, int sym, const T* r,
                    const T* p_old, T* p_new, T* Ap, int k, double* partials, Scalars* sc, hipStream_t st);
template <class T>
void lanczos_step_b(const Geom& G, const T* Ap, const T* dg, int sym, T* r, int k, double* partials, Scalars* sc,
                    hipStream_t st);
// mode 0 (SHK_CHEB_BOUNDS_GERSHGORIN): lambda_min from the estimate in the
// D-inner product, lambda_max the Gershgorin bound of D^-1 A (gershgorin() first);
// mode 1 (SHK_CHEB_BOUNDS_REFERENCE): both from the estimate, as the reference.
void cheb_bounds(int nsteps, int degree, int use_estimate, int mode, Scalars* sc, hipStream_t st);
template <class T>
void gershgorin(const Geom& G, const T* dg, const T* aE, const T* aN, double* partials, Scalars* sc,
                hipStream_t st);
// ||b - A h||_2 and ||b||_2 of the FP64 system for h = x + x0 (x0 may be
// null) into out[0], out[1] (device memory); partials: 2 x G.nblocks doubles
// ||b - A(x + x0)||^2 and ||b||^2 into out[0..1] (x0 may be null); with
// h_out, also h_out = x + x0 (under-relaxed towards x0 when relax > 0)
// ppart (with h_out and sc): also the Picard test of h_out against x0, the
// flag into sc (partials of the same count as the norms').
// The workgroups (partials) residual_norms launches for a grid.
int residual_norms_nblocks(const Geom& G);
void residual_norms(const Geom& G, const double* aP, const double* aE, const double* aN, const double* b,
                    const double* x, const double* x0, double* partials, double* out, hipStream_t st,
                    double* h_out = nullptr, double relax = 0.0, double* ppart = nullptr, Scalars* sc = nullptr);
// The fields a solve starts from (done, iter, tolerances, ...), reset on the device.
void reset_scalars(Scalars* sc, double atol, double rtol, int itmax, hipStream_t st);

// Preconditioned CG pieces (mode 0: set-up, mode 1: iteration).
template <class T>
void cg_init(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, const T* b, const T* x0, T* r, T* z,
             T* xacc, T* p_zero, int jacobi, double* partials, Scalars* sc, hipStream_t st);
template <class T>
void cheb_pass(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, const T* r, const T* s_in,
               const T* pc_in, T* s_out, T* pc_out, T* y, int k, int degree, int mode, double* partials, Scalars* sc,
               hipStream_t st);
template <class T>
void cheb_degree1(const Geom& G, const T* dg, const T* r, T* y, int mode, double* partials, Scalars* sc,
                  hipStream_t st);
template <class T>
void cg_matvec(const Geom& G, const T* aP, const T* aE, const T* aN, const T* z, const T* p_old, T* p_new, T* Ap,
               double* partials, Scalars* sc, hipStream_t st);
// FP64 CG with the finishing kernels folded into the next kernel: the matvec
// sums the pending r'z partials (pa), writes p'Ap partials to pb and returns
// their count, and takes the previous iteration's x += alpha p_old; the
// update sums both and takes r -= alpha Ap.
// The stop decision at the end of a batch of deferred iterations (done_defer
// and the pending x += alpha p_old when the solve stops there).  xs: the
// x-step mode of the matvec it stands in for (1 single steps; 0, 2 the odd and
// even iterations of paired steps, 2 also taking the owed step from p_pend).
void cg_defer_stop(size_t n, Scalars* sc, const double* pa, int nba, const double* p_old, double* x,
                   hipStream_t st, int xs = 1, const double* p_pend = nullptr);
// The owed x step of paired steps that no later kernel took (end of a solve).
void cg_xflush(size_t n, const Scalars* sc, double* x, hipStream_t st);
int cg_matvec_deferred(const Geom& G, const double* aP, const double* aE, const double* aN, const double* z,
                       const double* p_old, double* p_new, double* Ap, double* pb, Scalars* sc, const double* pa,
                       int nba, double* x, hipStream_t st);
int cg_matvec_deferred(const Geom& G, const double* aP, const double* aE, const double* aN, const float* z,
                       const double* p_old, double* p_new, double* Ap, double* pb, Scalars* sc, const double* pa,
                       int nba, double* x, hipStream_t st, int xs = 1);
void cg_update_deferred(const Geom& G, const double* Ap, double* r, Scalars* sc, const double* pa, int nba,
                        const double* pb, int nbb, hipStream_t st);
int mg_up_nblocks(int nx, int ny, int rb, int nu = 1);
template <class T>
void cg_update(const Geom& G, const T* dg, const T* p, const T* Ap, T* x, T* r, T* z, int jacobi, double* partials,
               Scalars* sc, hipStream_t st);
// Fused CG update + Chebyshev application (2 <= degree <= 8); r is ping-ponged.
// Returns -1 for an unsupported degree.
template <class T>
int cg_update_cheb(const Geom& G, int degree, const T* aP, const T* aE, const T* aN, const T* dg, const T* p,
                   const T* Ap, const T* r_old, T* r_new, T* x, T* z, double* partials, Scalars* sc,
                   hipStream_t st);
void cg_finish(int n, const double* x, const double* h_prev, double* h_new, double relax, hipStream_t st);

// Geometric multigrid on FP32 flux-form levels (see shk_kernels.hip).
// dot_mode: -1 no dot, 0 CG set-up gamma, 1 CG iteration gamma.
void mg_coarsen(int nxf, int nyf, int nxc, int nyc, const float* cf, const float* ef, const float* nf, float* cc,
                float* ec, float* nc, hipStream_t st);
void mg_diag(int nx, int ny, const float* c, const float* e, const float* n, float* D, hipStream_t st);
// the linear transfers: the lumped Galerkin coarse operator, the restriction
// of a level's residual from its stored x (FP32 or BF16 coefficients), and
// u = x + scale P xc in place
void mg_coarsen_lin(int nxf, int nyf, int nxc, int nyc, const float* cf, const float* ef, const float* nf, float* cc,
                    float* ec, float* nc, hipStream_t st);
template <class TB, class CT>
void mg_restrict_lin(int nx, int ny, int nxc, int nyc, const CT* c, const CT* e, const CT* n, const TB* b,
                     const float* x, float* bc, const Scalars* sc, hipStream_t st);
void mg_prolong_lin(int nx, int ny, int nxc, float scale, float* x, const float* xc, const Scalars* sc,
                    hipStream_t st);
// the red-black legs with the linear transfers fused in, where the level's
// legs are the two-row kernels (nu sweeps each way where mg_lin_nu2 allows
// two); false (nothing launched) elsewhere
template <class TB, class CT>
bool mg_down_lin(int nx, int ny, int nxc, float omega, const CT* c, const CT* e, const CT* n, const CT* D,
                 const TB* b, float* x, float* bc, const Scalars* sc, int nu, hipStream_t st);
template <class TB, class CT, class TO>
bool mg_up_lin(int nx, int ny, int nxc, float omega, float scale, const CT* c, const CT* e, const CT* n,
               const CT* D, const TB* b, const float* x, const float* xc, TO* out, int dot_mode, double* partials,
               Scalars* sc, int nu, hipStream_t st);
// whether an nx x ny level with the linear transfers takes two sweeps each way
// (nu = 2 in mg_down_lin / mg_up_lin)
bool mg_lin_nu2(int nx, int ny);
template <class TB>
void mg_down(int nx, int ny, int nxc, int nyc, float omega, const float* c, const float* e, const float* n,
             const float* D, const TB* b, float* x, float* bc, const Scalars* sc, int rb, hipStream_t st);
template <class TB>
void mg_up(int nx, int ny, int nxc, float omega, float scale, const float* c, const float* e, const float* n,
           const float* D, const TB* b, const float* x, const float* xc, TB* out, int dot_mode,
           double* partials, Scalars* sc, int rb, hipStream_t st);
void mg_tail(const MGTail& T, const Scalars* sc, hipStream_t st);
bool mg_bf16_ok(int nx, int ny);
void to_bf16x4(size_t n, const float* a0, const float* a1, const float* a2, const float* a3, uint16_t* o0,
               uint16_t* o1, uint16_t* o2, uint16_t* o3, hipStream_t st);
// The finest level's BF16 legs; nu = 2: two red-black sweeps each way
// (k_mg_down_sw2 / k_mg_up_sw2), one otherwise.
template <class TB>
void mg_down_bf16(int nx, int ny, int nxc, float omega, const uint16_t* c, const uint16_t* e, const uint16_t* n,
                  const uint16_t* D, const TB* b, float* x, float* bc, const Scalars* sc, int nu, hipStream_t st);
// The CG update r_out = r - alpha Ap fused into the finest down-leg (FP64 CG
// with the deferred finish): alpha from the pending p'Ap (pb, nbb) and r'z
// (pa, nba) partials, stored to sc by the kernel as the update would.
struct CGUpd {
    const double* Ap = nullptr;
    double* b_out = nullptr;
    Scalars* sc = nullptr;
    const double* pa = nullptr;
    int nba = 0;
    const double* pb = nullptr;
    int nbb = 0;
};
// Whether the finest level of an nx x ny grid takes the leg that can fuse the
// update (the BF16 marching leg at more than two rows per wave).
bool mg_down_fuses_update(int nx, int ny);
void mg_down_bf16_upd(int nx, int ny, int nxc, float omega, const uint16_t* c, const uint16_t* e, const uint16_t* n,
                      const uint16_t* D, const double* b, float* x, float* bc, const CGUpd& U, int nu,
                      hipStream_t st);
template <class TB, class TO>
void mg_up_bf16(int nx, int ny, int nxc, float omega, float scale, const uint16_t* c, const uint16_t* e,
                const uint16_t* n, const uint16_t* D, const TB* b, const float* x, const float* xc, TO* out,
                int dot_mode, double* partials, Scalars* sc, int nu, hipStream_t st);
// dynamic LDS bytes of the tail kernel (rhs, x, c, e, n, D per cell and the
// coarsest inverse), and the opt-in for more than the default 64 KB (0 = ok)
size_t mg_tail_stride(size_t tail_cells);
size_t mg_tail_lds(size_t tail_cells, size_t coarsest_cells, size_t hcells = 0);
int mg_tail_reserve(size_t bytes);
void mg_dense_inverse(int nx, int ny, const float* c, const float* e, const float* n, float* inv, hipStream_t st);

// Mixed precision: FP32 copies of the coefficients; the FP64 residual
// r = b - A x with its Jacobi-weighted norm (first != 0 records the initial
// norm) and its FP32 copy; the FP64 correction x += d.
void to_f32(size_t n, const double* a, float* b, hipStream_t st);
void read_flush(size_t n, const double* a, double* sink, hipStream_t st);
// learned warm start (network input / output on the image padded to W8 x H8)
void nn_resid_scale(size_t n, const double* aP, const double* rhs, const double* Ap, double* r, double* partials,
                    Scalars* sc, hipStream_t st);
// slot 0: the warm start's scale and step; slot 1: the preconditioner's
void nn_features(int nx, int ny, int W8, int H8, const double* aP, const double* aE, const double* aN,
                 const double* cP, const double* r, const Scalars* sc, int slot, void* x, hipStream_t st);
void nn_correction(int nx, int ny, int W8, const float* N, const double* aP, const double* aE, const double* aN,
                   const Scalars* sc, int slot, double* e, hipStream_t st);
void nn_line_search(size_t n, const double* e, const double* r, const double* Ae, double* partials, Scalars* sc,
                    int slot, hipStream_t st);
void nn_warm(size_t n, const double* h, const double* e, const Scalars* sc, double* x0, hipStream_t st);
void nn_target(int nx, int ny, int W8, int H8, const double* aP, const double* h_new, const double* h_prev,
               const Scalars* sc, float* t, hipStream_t st);
void nn_ref(Scalars* sc, int mode, double n, hipStream_t st);
// learned preconditioner inside flexible CG
void nn_target_cg(int nx, int ny, int W8, int H8, const double* aP, const double* x_final, const double* x_k,
                  const double* s_k, float* t, hipStream_t st);
void nn_rscale(size_t n, const double* aP, const double* r, double* partials, Scalars* sc, hipStream_t st);
void nn_r1(size_t n, const double* r, const double* Ae, const Scalars* sc, double* r1, hipStream_t st);
void nn_combine(size_t n, const double* r, const double* e, double* z, double* partials, Scalars* sc,
                hipStream_t st);
void fcg_beta(size_t n, const double* z, const double* Ap, double* partials, Scalars* sc, hipStream_t st);
void ir_residual(const Geom& G, const double* aP, const double* aE, const double* aN, const double* b,
                 const double* x, float* r32, int first, double* partials, Scalars* sc, hipStream_t st);
void ir_correct(size_t n, double* x, const float* d, hipStream_t st);
void relax(size_t n, double* h, const double* h_prev, double alpha, hipStream_t st);
}  // namespace launch

}  // namespace shk

// The synthetic glacier generator (shk_gen.hip): writes the raw inputs of
// set_initial_conditions into F, and the moulin field when moulin is given,
// using w and tmp (nx*ny each) as scratch.  Synchronous on st.
namespace shkgen {
int generate(const shk_grid& g, const shk_glacier_spec& sp, shk::Fields F, double* w, double* tmp, double* moulin,
             hipStream_t st, std::string& err);
}  // namespace shkgen
