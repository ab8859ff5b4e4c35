/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C FP64 restatement of Shakti.jl's hydraulic-head path, used by the
 * test-suite, by the bench's CPU-baseline leg and by the smoke check as the
 * checker of the HIP library.  Nothing in the product (the HIP C-ABI library
 * and its host mirror) links, loads or calls this file.
 *
 * Every routine cites the reference function it restates (paths relative to
 * the Shakti.jl repository root).  Floating-point expressions keep Julia's
 * left-to-right evaluation order and are compiled with -ffp-contract=off so no
 * multiply-add is fused where Julia would not fuse it.
 *
 * Layout (reference src/state.jl:13-57, src/grid.jl:26-35): column-major with
 * ix fastest.  Centre field (nx, ny): idx = ix + iy*nx.  x-face field
 * (nx+1, ny): idx = ix + iy*(nx+1).  y-face field (nx, ny+1): idx = ix + iy*nx.
 * All indices below are 0-based; the reference's 1-based `ix` is `ix+1` here.
 *
 * Reductions (dot products) are summed over a fixed set of chunks whose
 * boundaries do not depend on the thread count, so results are bit-identical
 * for any OMP_NUM_THREADS.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ types */

typedef struct { int32_t nx, ny; double lx, ly, dx, dy, dx2, dy2; } orc_grid;

/* src/model_parameters.jl:9-29, defaults :55-84 */
typedef struct {
    double rho_w, rho_sw, rho_i, g, nu, n, omega, L, br, lr, ct, cw, p_atm, b_min, b_max, e_v;
} orc_params;

/* The scheme choices a Simulation bundles (src/simulation.jl:105-146). */
typedef struct {
    int32_t kfs;          /* 0 Arithmetic, 1 Harmonic (src/k_face_scheme.jl:15,23) */
    int32_t gap_scheme;   /* 0 implicit, 1 explicit (src/simulation.jl:122-128) */
    int32_t sliding_law;  /* 0 RegularizedCoulomb, 1 Prescribed, 2 Linear, 3 RegularizedCoulombField,
                             4 RegularizedCoulombV0 (src/sliding_law.jl:33, 45, 61, 270, 422) */
    double sliding_C;     /* RegularizedCoulombSlidingLaw.C */
    double v0;            /* RegularizedCoulombV0SlidingLaw.v0 */
    const double* Cx;     /* laws 2-4: C staggered onto x-faces ((nx+1) x ny; law 2 squares it at use) */
    const double* Cy;     /*           and y-faces (nx x (ny+1)) by orc_stagger_c */
} orc_model;

/* Every State field (src/state.jl:13-57), in the reference's order. */
typedef struct {
    double *h, *pw, *po, *b, *beta, *abs_ub, *mdot, *shear, *potential, *sensible, *Re, *K, *G, *zb,
        *zs, *H, *ieb, *lambda, *A_visc, *N, *mask;
    double *dhdx, *q_x, *Re_x, *b_x, *ub_x, *taub_x, *dpwdx, *valid_x;
    double *dhdy, *q_y, *Re_y, *b_y, *ub_y, *taub_y, *dpwdy, *valid_y;
} orc_state;

/* src/mask.jl:21-41 */
#define GROUNDED 0.0
#define OCEAN 1.0
#define LAND 2.0
#define OTHER_BASIN 3.0

#define EPS64 2.220446049250313e-16 /* eps(Float64) */

#define C_(ix, iy) ((size_t)(ix) + (size_t)(iy) * (size_t)nx)
#define XF_(ix, iy) ((size_t)(ix) + (size_t)(iy) * (size_t)(nx + 1))
#define YF_(ix, iy) ((size_t)(ix) + (size_t)(iy) * (size_t)nx)

/* ------------------------------------------------------------ scalar math */

/* x^k for an integer exponent, as Julia's Base.:^(::Float64, ::Integer) does
 * it: power-by-squaring with an error-compensated product (pow_body), k==3
 * special-cased to x*x*x.  Shakti's `pow` (src/model_parameters.jl:133-134)
 * sends integer-valued exponents here via canonical_exponent (:122-123). */
static double two_prod_err(double a, double b, double p) { return fma(a, b, -p); }

static double jl_pow_int(double x, long n) {
    if (n == 0) return 1.0;
    if (n == 3) return x * x * x;
    if (n < 0) return 1.0 / jl_pow_int(x, -n); /* not reached for Glen exponents >= 1 */
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = y * xnlo + x * ynlo; /* muladd(y, xnlo, x*ynlo) */
            double yy = x * y;
            ynlo = two_prod_err(x, y, yy);
            y = yy;
            ynlo += err;
        }
        double err = x * 2.0 * xnlo;
        double xx = x * x;
        xnlo = two_prod_err(x, x, xx);
        x = xx;
        xnlo += err;
        n >>= 1;
    }
    double err = y * xnlo + x * ynlo;
    if (isfinite(x) && isfinite(err)) return fma(x, y, err);
    return x * y;
}

/* pow(x, e) with e already canonicalised: integer-valued e takes the integer
 * path, anything else the general floating-point power. */
static double jl_pow(double x, double e) {
    if (e == floor(e) && fabs(e) < 1e9) return jl_pow_int(x, (long)e);
    return pow(x, e);
}

/* Julia max(0, v) / clamp semantics with NaN propagation. */
static double jl_max0(double v) { return isnan(v) ? v : (v > 0.0 ? v : 0.0); }
static double jl_clamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }
static double jl_min(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b; /* min(-0.0, 0.0) == -0.0 */
}
/* Julia sign(x): 0 -> x itself (keeps the sign of zero), NaN -> NaN, else copysign(1, x). */
static double jl_sign(double x) { return (x == 0.0 || isnan(x)) ? x : copysign(1.0, x); }
/* Julia Bool * Float: true*x == x, false*x == copysign(0, x). */
static double jl_boolmul(int t, double x) { return t ? x : copysign(0.0, x); }

/* ------------------------------------------------------- grid & parameters */

/* src/grid.jl:26-35 */
void orc_make_grid(int32_t nx, int32_t ny, double lx, double ly, orc_grid* g) {
    g->nx = nx; g->ny = ny; g->lx = lx; g->ly = ly;
    g->dx = lx / (double)(nx - 1);
    g->dy = ly / (double)(ny - 1);
    g->dx2 = g->dx * g->dx;
    g->dy2 = g->dy * g->dy;
}

/* src/model_parameters.jl:55-84 defaults */
void orc_default_params(orc_params* p) {
    p->rho_w = 1000.0; p->rho_sw = 1027.0; p->rho_i = 910.0; p->g = 9.81; p->nu = 1.787e-6;
    p->n = 3.0; p->omega = 1e-4; p->L = 334e3; p->br = 0.05; p->lr = 2.0; p->ct = 7.5e-8;
    p->cw = 4.22e3; p->p_atm = 0.0; p->b_min = 0.0; p->b_max = INFINITY; p->e_v = 0.0;
}

/* ------------------------------------------------------------- mask logic */

/* src/mask.jl:62-104 compute_valid_x/y_kernel!, compute_face_masks! */
void orc_face_masks(const orc_grid* g, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            double v = 1.0;
            if (ix > 0 && ix < nx)
                v = (s->mask[C_(ix - 1, iy)] != OTHER_BASIN && s->mask[C_(ix, iy)] != OTHER_BASIN) ? 1.0 : 0.0;
            s->valid_x[XF_(ix, iy)] = v;
        }
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            double v = 1.0;
            if (iy > 0 && iy < ny)
                v = (s->mask[C_(ix, iy - 1)] != OTHER_BASIN && s->mask[C_(ix, iy)] != OTHER_BASIN) ? 1.0 : 0.0;
            s->valid_y[YF_(ix, iy)] = v;
        }
}

/* src/mask.jl:117-150 apply_mask_to_sliding_x/y_kernel! */
void orc_apply_mask_to_sliding(const orc_grid* g, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix)
            if (ix == 0 || ix == nx ||
                !(s->mask[C_(ix - 1, iy)] == GROUNDED && s->mask[C_(ix, iy)] == GROUNDED))
                s->ub_x[XF_(ix, iy)] = 0.0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            if (iy == 0 || iy == ny ||
                !(s->mask[C_(ix, iy - 1)] == GROUNDED && s->mask[C_(ix, iy)] == GROUNDED))
                s->ub_y[YF_(ix, iy)] = 0.0;
}

/* ------------------------------------------------- per-cell physics pieces */

/* src/field_gradients.jl:16-127 (interior faces only; boundary faces untouched) */
static void grad_xy(const orc_grid* g, const double* f, const double* valid_x, const double* valid_y,
                    double* dfdx, double* dfdy) {
    const int nx = g->nx, ny = g->ny;
    const double dx = g->dx, dy = g->dy;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 1; ix < nx; ++ix)
            dfdx[XF_(ix, iy)] = ((f[C_(ix, iy)] - f[C_(ix - 1, iy)]) / dx) * valid_x[XF_(ix, iy)];
#pragma omp parallel for schedule(static)
    for (int iy = 1; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            dfdy[YF_(ix, iy)] = ((f[C_(ix, iy)] - f[C_(ix, iy - 1)]) / dy) * valid_y[YF_(ix, iy)];
}

/* src/pressure.jl:1-13 */
static void compute_pw(const orc_grid* g, const orc_params* p, orc_state* s) {
    const size_t n = (size_t)g->nx * g->ny;
    const double rg = p->rho_w * p->g;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) s->pw[i] = rg * (s->h[i] - s->zb[i]);
}

/* src/pressure.jl:15-33 */
static void compute_N(const orc_grid* g, orc_state* s) {
    const size_t n = (size_t)g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) s->N[i] = jl_boolmul(s->mask[i] == GROUNDED, s->po[i] - s->pw[i]);
}

/* src/water_flux.jl:192-216 compute_q_and_Re_xy_kernel! (exact, lag-free) */
static void qre_face(double bf, double grad, const orc_params* p, double* q, double* Re) {
    double G = fabs(grad);
    double D = (bf * bf * bf * p->g * G) / (12.0 * (p->nu * p->nu));
    double R = 2.0 * D / (1.0 + sqrt(1.0 + 4.0 * p->omega * D));
    *Re = R;
    *q = -jl_sign(grad) * p->nu * R;
}
static void compute_q_and_Re_xy(const orc_grid* g, const orc_params* p, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            size_t k = XF_(ix, iy);
            qre_face(s->b_x[k], s->dhdx[k], p, &s->q_x[k], &s->Re_x[k]);
        }
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = YF_(ix, iy);
            qre_face(s->b_y[k], s->dhdy[k], p, &s->q_y[k], &s->Re_y[k]);
        }
}

/* src/water_flux.jl:5-105 lagged pair (IC seed and the reference's own tests):
 * q from the previous Re, then Re = |q|/nu. */
void orc_lagged_q_then_Re(const orc_grid* g, const orc_params* p, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            size_t k = XF_(ix, iy);
            double bx = s->b_x[k];
            s->q_x[k] = -(bx * bx * bx * p->g * s->dhdx[k]) / (12.0 * p->nu * (1.0 + p->omega * s->Re_x[k]));
        }
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t k = YF_(ix, iy);
            double by = s->b_y[k];
            s->q_y[k] = -(by * by * by * p->g * s->dhdy[k]) / (12.0 * p->nu * (1.0 + p->omega * s->Re_y[k]));
        }
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) s->Re_x[XF_(ix, iy)] = fabs(s->q_x[XF_(ix, iy)]) / p->nu;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) s->Re_y[YF_(ix, iy)] = fabs(s->q_y[YF_(ix, iy)]) / p->nu;
}

/* src/water_flux.jl:107-119 */
static void compute_Re(const orc_grid* g, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            s->Re[C_(ix, iy)] = (s->Re_x[XF_(ix, iy)] + s->Re_x[XF_(ix + 1, iy)] + s->Re_y[YF_(ix, iy)] +
                                 s->Re_y[YF_(ix, iy + 1)]) / 4.0;
}

/* The constructors LinearSlidingLaw(g, C) (src/sliding_law.jl:73-90),
 * RegularizedCoulombFieldSlidingLaw(g, C) (:284-301) and
 * RegularizedCoulombV0SlidingLaw(g, C, v0) (:436-452): C at the cell centres
 * staggered onto the faces, outer faces the cell's value, inner faces the
 * mean of the two cells. */
void orc_stagger_c(const orc_grid* g, const double* C, double* Cx, double* Cy) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix)
            Cx[XF_(ix, iy)] = ix == 0 ? C[C_(0, iy)]
                              : ix == nx ? C[C_(nx - 1, iy)]
                                         : (C[C_(ix - 1, iy)] + C[C_(ix, iy)]) / 2.0;
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            Cy[YF_(ix, iy)] = iy == 0 ? C[C_(ix, 0)]
                              : iy == ny ? C[C_(ix, ny - 1)]
                                         : (C[C_(ix, iy - 1)] + C[C_(ix, iy)]) / 2.0;
}

/* The face stress of each law:
 * src/sliding_law.jl:213-256 compute_taub_xy_kernel! (RegularizedCoulomb, scalar C),
 * :360-396 compute_taub_xy_field_kernel! (the same with the face C),
 * :509-543 compute_taub_xy_v0_kernel! (v0 in place of |N|^n lambda),
 * :590-602 compute_taub_xy_linear_kernel! (Cx2 * Nf * ub, Cx2 the squared face C). */
static double taub_face(int law, double Nf, double abs_v, double lf, double ub, double C, double Cf, double v0,
                        double n, double inv_n) {
    if (law == 2) return (Cf * Cf) * Nf * ub;
    if (law == 4) return abs_v > 0.0 ? Nf * Cf * jl_pow(abs_v / (abs_v + v0), inv_n) * (ub / abs_v) : 0.0;
    const double c = law == 3 ? Cf : C;
    if (abs_v > 0.0)
        return Nf * c * jl_pow(abs_v / (abs_v + jl_pow(fabs(Nf), n) * lf), inv_n) * (ub / abs_v);
    return 0.0;
}
static void compute_taub_xy(const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s) {
    if (m->sliding_law == 1) return; /* Prescribed: no-op (src/sliding_law.jl:115-131) */
    const int nx = g->nx, ny = g->ny, law = m->sliding_law;
    const double C = m->sliding_C, n = p->n, inv_n = 1.0 / p->n, v0 = m->v0;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            double Nf, av, lf;
            if (ix == 0) {
                Nf = s->N[C_(0, iy)]; av = s->abs_ub[C_(0, iy)]; lf = s->lambda[C_(0, iy)];
            } else if (ix == nx) {
                Nf = s->N[C_(nx - 1, iy)]; av = s->abs_ub[C_(nx - 1, iy)]; lf = s->lambda[C_(nx - 1, iy)];
            } else {
                Nf = (s->N[C_(ix, iy)] + s->N[C_(ix - 1, iy)]) / 2.0;
                lf = (s->lambda[C_(ix, iy)] + s->lambda[C_(ix - 1, iy)]) / 2.0;
                av = (s->abs_ub[C_(ix, iy)] + s->abs_ub[C_(ix - 1, iy)]) / 2.0;
            }
            s->taub_x[XF_(ix, iy)] = taub_face(law, Nf, av, lf, s->ub_x[XF_(ix, iy)], C, law >= 2 ? m->Cx[XF_(ix, iy)] : 0.0,
                                               v0, n, inv_n);
        }
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            double Nf, av, lf;
            if (iy == 0) {
                Nf = s->N[C_(ix, 0)]; av = s->abs_ub[C_(ix, 0)]; lf = s->lambda[C_(ix, 0)];
            } else if (iy == ny) {
                Nf = s->N[C_(ix, ny - 1)]; av = s->abs_ub[C_(ix, ny - 1)]; lf = s->lambda[C_(ix, ny - 1)];
            } else {
                Nf = (s->N[C_(ix, iy)] + s->N[C_(ix, iy - 1)]) / 2.0;
                lf = (s->lambda[C_(ix, iy)] + s->lambda[C_(ix, iy - 1)]) / 2.0;
                av = (s->abs_ub[C_(ix, iy)] + s->abs_ub[C_(ix, iy - 1)]) / 2.0;
            }
            s->taub_y[YF_(ix, iy)] = taub_face(law, Nf, av, lf, s->ub_y[YF_(ix, iy)], C, law >= 2 ? m->Cy[YF_(ix, iy)] : 0.0,
                                               v0, n, inv_n);
        }
}

/* src/melt_rate.jl:106-168 compute_mdot_kernel! (both methods); sensible heat is
 * on iff ct*cw != 0 (src/simulation.jl:132). */
static void compute_mdot(const orc_grid* g, const orc_params* p, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    const int with_sens = !(p->ct == 0.0 || p->cw == 0.0);
    const double Linv = 1.0 / p->L, rho_w = p->rho_w, gg = p->g, ct = p->ct, cw = p->cw;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t e = XF_(ix + 1, iy), w = XF_(ix, iy), nn = YF_(ix, iy + 1), ss = YF_(ix, iy), c = C_(ix, iy);
            double sh = fabs((s->ub_x[e] * s->taub_x[e] + s->ub_x[w] * s->taub_x[w]) / 2.0 +
                             (s->ub_y[nn] * s->taub_y[nn] + s->ub_y[ss] * s->taub_y[ss]) / 2.0);
            double pot = fabs((s->q_x[e] * s->dhdx[e] + s->q_x[w] * s->dhdx[w]) / 2.0 +
                              (s->q_y[nn] * s->dhdy[nn] + s->q_y[ss] * s->dhdy[ss]) / 2.0);
            s->shear[c] = sh;
            s->potential[c] = pot;
            if (with_sens) {
                double sens = (s->q_x[e] * s->dpwdx[e] + s->q_x[w] * s->dpwdx[w]) / 2.0 +
                              (s->q_y[nn] * s->dpwdy[nn] + s->q_y[ss] * s->dpwdy[ss]) / 2.0;
                s->sensible[c] = sens;
                s->mdot[c] = Linv * (s->G[c] + sh + rho_w * gg * pot + ct * cw * rho_w * sens);
            } else {
                s->mdot[c] = Linv * (s->G[c] + sh + rho_w * gg * pot);
            }
        }
}

/* src/water_flux.jl:221-234 */
static void compute_K(const orc_grid* g, const orc_params* p, orc_state* s) {
    const size_t n = (size_t)g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        double b = s->b[i];
        s->K[i] = (b * b * b * p->g) / (12.0 * p->nu * (1.0 + p->omega * s->Re[i]));
    }
}

/* src/gap_height.jl:47-73; skipped entirely when br == 0 (src/simulation.jl:136) */
static void compute_beta(const orc_grid* g, const orc_params* p, orc_state* s) {
    if (p->br == 0.0) return;
    const size_t n = (size_t)g->nx * g->ny;
    for (size_t i = 0; i < n; ++i) s->beta[i] = jl_max0((p->br - s->b[i]) / p->lr);
}

/* src/gap_height.jl:75-115 */
static void compute_b_xy(const orc_grid* g, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            double v;
            if (ix == 0) v = s->b[C_(0, iy)];
            else if (ix == nx) v = s->b[C_(nx - 1, iy)];
            else v = (s->b[C_(ix, iy)] + s->b[C_(ix - 1, iy)]) / 2.0;
            s->b_x[XF_(ix, iy)] = v;
        }
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            double v;
            if (iy == 0) v = s->b[C_(ix, 0)];
            else if (iy == ny) v = s->b[C_(ix, ny - 1)];
            else v = (s->b[C_(ix, iy)] + s->b[C_(ix, iy - 1)]) / 2.0;
            s->b_y[YF_(ix, iy)] = v;
        }
}

/* src/static_fields.jl:59-74 */
static void compute_abs_ub(const orc_grid* g, orc_state* s) {
    const int nx = g->nx, ny = g->ny;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            double ubx = (s->ub_x[XF_(ix, iy)] + s->ub_x[XF_(ix + 1, iy)]) / 2.0;
            double uby = (s->ub_y[YF_(ix, iy)] + s->ub_y[YF_(ix, iy + 1)]) / 2.0;
            s->abs_ub[C_(ix, iy)] = sqrt(ubx * ubx + uby * uby);
        }
}

/* ------------------------------------------------------ init and the tail */

/* src/initial_conditions.jl:14-126.  The raw inputs (mask, A_visc, zb, zs, b,
 * G, ub_x, ub_y, ieb, taub_x, taub_y) must already be in `s`. */
void orc_set_initial_conditions(const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s) {
    const size_t n = (size_t)g->nx * g->ny;
    orc_face_masks(g, s);
    for (size_t i = 0; i < n; ++i) s->lambda[i] = 1.5 * s->A_visc[i];
    for (size_t i = 0; i < n; ++i)
        if (s->mask[i] != GROUNDED) s->b[i] = 0.0;
    for (size_t i = 0; i < n; ++i) s->H[i] = s->zs[i] - (s->zb[i] + s->b[i]); /* static_fields.jl:8-13 */
    compute_beta(g, p, s);
    orc_apply_mask_to_sliding(g, s);
    compute_abs_ub(g, s);
    for (size_t i = 0; i < n; ++i) s->po[i] = p->rho_i * p->g * s->H[i]; /* static_fields.jl:22-27 */
    for (size_t i = 0; i < n; ++i) {
        double m_ = s->mask[i];
        double pw = s->po[i] / 2.0;
        if (m_ == OCEAN) pw = p->p_atm - p->rho_sw * p->g * jl_min(s->zb[i], 0.0);
        else if (m_ == LAND || m_ == OTHER_BASIN) pw = p->p_atm;
        s->pw[i] = pw;
    }
    grad_xy(g, s->pw, s->valid_x, s->valid_y, s->dpwdx, s->dpwdy);
    compute_N(g, s);
    for (size_t i = 0; i < n; ++i) s->h[i] = s->pw[i] / (p->rho_w * p->g) + s->zb[i]; /* static_fields.jl:39-44 */
    grad_xy(g, s->h, s->valid_x, s->valid_y, s->dhdx, s->dhdy);
    compute_b_xy(g, s);
    compute_q_and_Re_xy(g, p, s);
    compute_Re(g, s);
    compute_taub_xy(g, p, m, s);
    compute_mdot(g, p, s);
    compute_K(g, p, s);
}

/* The field refresh after every linear solve, in the reference's fixed order:
 * src/elliptic_solver.jl:193-206 (Picard_iteration!) and
 * src/parabolic_solver.jl:23-36 (parabolic_solver!). */
void orc_update_derived(const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s) {
    grad_xy(g, s->h, s->valid_x, s->valid_y, s->dhdx, s->dhdy);
    compute_pw(g, p, s);
    grad_xy(g, s->pw, s->valid_x, s->valid_y, s->dpwdx, s->dpwdy);
    compute_N(g, s);
    compute_q_and_Re_xy(g, p, s);
    compute_Re(g, s);
    compute_taub_xy(g, p, m, s);
    compute_mdot(g, p, s);
    compute_K(g, p, s);
}

/* The mid-solve perturbation sequence of the reference's assembly tests
 * (test/linear_solver_test.jl:38-54): separate gradients, pw, N, the LAGGED
 * q/Re pair, Re, taub, mdot, K.  `h` must already be perturbed. */
void orc_test_refresh_lagged(const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s) {
    grad_xy(g, s->h, s->valid_x, s->valid_y, s->dhdx, s->dhdy);
    compute_pw(g, p, s);
    grad_xy(g, s->pw, s->valid_x, s->valid_y, s->dpwdx, s->dpwdy);
    compute_N(g, s);
    orc_lagged_q_then_Re(g, p, s);
    compute_Re(g, s);
    compute_taub_xy(g, p, m, s);
    compute_mdot(g, p, s);
    compute_K(g, p, s);
}

/* ---------------------------------------------------------------- gap step */

/* src/run.jl:123-164 (compute_b!, step_b!) with src/gap_height.jl:117-135 */
void orc_step_b(const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s, double dt) {
    const size_t n = (size_t)g->nx * g->ny;
    const double e = p->n - 1.0;
    for (size_t i = 0; i < n; ++i) {
        if (s->mask[i] != GROUNDED) continue;
        double Np = jl_pow(fabs(s->N[i]), e);
        double v;
        if (m->gap_scheme == 0)
            v = (s->b[i] + dt * (s->mdot[i] / p->rho_i + s->beta[i] * s->abs_ub[i])) /
                (1.0 + dt * s->A_visc[i] * Np * s->N[i]);
        else
            v = s->b[i] + dt * (s->mdot[i] / p->rho_i + s->beta[i] * s->abs_ub[i] - s->A_visc[i] * Np * s->N[i] * s->b[i]);
        s->b[i] = jl_clamp(v, p->b_min, p->b_max);
    }
    compute_beta(g, p, s);
    compute_b_xy(g, s);
}

/* ------------------------------------------------------------- assembly */

/* src/linear_solver.jl:13-27 */
static double dirichlet_head(double m, double zb, const orc_params* p) {
    if (m == OCEAN) {
        double pw_bc = p->p_atm - p->rho_sw * p->g * jl_min(zb, 0.0);
        return pw_bc / (p->rho_w * p->g) + zb;
    }
    return p->p_atm / (p->rho_w * p->g) + zb;
}

/* src/k_face_scheme.jl:32-61 */
static double boundary_K_face(int kfs, const double* K, const double* mask, size_t c1, size_t c2) {
    double m2 = mask[c2];
    if (m2 == OTHER_BASIN) return 0.0;
    if (m2 == OCEAN || m2 == LAND) return K[c1];
    if (kfs == 0) return (K[c1] + K[c2]) / 2.0;
    return (2.0 * K[c1] * K[c2]) / (K[c1] + K[c2] + EPS64);
}

/* src/linear_solver.jl:438-523 update_MFLS_kernel!/update_MFLS! (parabolic == 0)
 * and :529-610 update_MFLS_parabolic_kernel!/update_MFLS_parabolic! (parabolic == 1).
 * Off-diagonals are stored positive; the matvec applies the minus sign. */
void orc_assemble_mfls(const orc_grid* g, const orc_params* p, int32_t kfs, const orc_state* s, int32_t parabolic,
                       double dt, double* aP, double* aE, double* aW, double* aN, double* aS, double* rhs) {
    const int nx = g->nx, ny = g->ny;
    const size_t n = (size_t)nx * ny;
    memset(aP, 0, n * sizeof(double)); memset(aE, 0, n * sizeof(double));
    memset(aW, 0, n * sizeof(double)); memset(aN, 0, n * sizeof(double));
    memset(aS, 0, n * sizeof(double)); memset(rhs, 0, n * sizeof(double));
    const double e = p->n - 1.0, rg = p->rho_w * p->g;
    const double *mask = s->mask, *K = s->K, *zb = s->zb;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t c = C_(ix, iy);
            double m = mask[c];
            if (m == OCEAN || m == LAND) { aP[c] = 1.0; rhs[c] = dirichlet_head(m, zb[c], p); continue; }
            if (m == OTHER_BASIN) { aP[c] = 1.0; rhs[c] = s->h[c]; continue; }
            double fE = ix < nx - 1 ? boundary_K_face(kfs, K, mask, c, C_(ix + 1, iy)) / g->dx2 : 0.0;
            double fW = ix > 0 ? boundary_K_face(kfs, K, mask, c, C_(ix - 1, iy)) / g->dx2 : 0.0;
            double fN = iy < ny - 1 ? boundary_K_face(kfs, K, mask, c, C_(ix, iy + 1)) / g->dy2 : 0.0;
            double fS = iy > 0 ? boundary_K_face(kfs, K, mask, c, C_(ix, iy - 1)) / g->dy2 : 0.0;
            double Np = jl_pow(fabs(s->N[c]), e);
            if (!parabolic)
                aP[c] = (fE + fW + fN + fS) + p->n * p->rho_w * p->g * s->A_visc[c] * Np * s->b[c];
            else
                aP[c] = (fE + fW + fN + fS) + p->e_v / dt;
            double drhs = 0.0;
            if (ix < nx - 1) {
                size_t o = C_(ix + 1, iy); double mo = mask[o];
                if (mo == OCEAN || mo == LAND) drhs += fE * dirichlet_head(mo, zb[o], p); else aE[c] = fE;
            }
            if (ix > 0) {
                size_t o = C_(ix - 1, iy); double mo = mask[o];
                if (mo == OCEAN || mo == LAND) drhs += fW * dirichlet_head(mo, zb[o], p); else aW[c] = fW;
            }
            if (iy < ny - 1) {
                size_t o = C_(ix, iy + 1); double mo = mask[o];
                if (mo == OCEAN || mo == LAND) drhs += fN * dirichlet_head(mo, zb[o], p); else aN[c] = fN;
            }
            if (iy > 0) {
                size_t o = C_(ix, iy - 1); double mo = mask[o];
                if (mo == OCEAN || mo == LAND) drhs += fS * dirichlet_head(mo, zb[o], p); else aS[c] = fS;
            }
            if (!parabolic)
                rhs[c] = s->mdot[c] * (1.0 / p->rho_w - 1.0 / p->rho_i) - s->beta[c] * s->abs_ub[c] +
                         s->A_visc[c] * Np * s->N[c] * s->b[c] +
                         s->A_visc[c] * Np * (p->n * rg * s->h[c]) * s->b[c] + s->ieb[c] + drhs;
            else
                rhs[c] = s->mdot[c] * (1.0 / p->rho_w - 1.0 / p->rho_i) - s->beta[c] * s->abs_ub[c] +
                         s->A_visc[c] * Np * s->N[c] * s->b[c] + s->ieb[c] + (p->e_v / dt) * s->h[c] + drhs;
        }
    (void)rg;
}

/* src/linear_solver.jl:144-198 + :226-333 (update_SALS_kernel!/update_SALS!)
 * and :345-430 (parabolic).  The sparse matrix is produced as CSR rows of the
 * 5-point pattern in the reference's push order per row (diag, W, E, S, N);
 * `cols` and `vals` hold 5 slots per row, unused slots have col = -1.
 * Off-diagonal entries carry the minus sign, as in nzval. */
void orc_assemble_sals(const orc_grid* g, const orc_params* p, int32_t kfs, const orc_state* s, int32_t parabolic,
                       double dt, int64_t* cols, double* vals, double* rhs) {
    const int nx = g->nx, ny = g->ny;
    const size_t n = (size_t)nx * ny;
    const double e = p->n - 1.0;
    for (size_t r = 0; r < n; ++r) {
        for (int k = 0; k < 5; ++k) { cols[5 * r + k] = -1; vals[5 * r + k] = 0.0; }
        rhs[r] = 0.0;
    }
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t row = C_(ix, iy);
            int64_t* cr = cols + 5 * row;
            double* vr = vals + 5 * row;
            /* pattern: slot 0 diag, 1 W, 2 E, 3 S, 4 N (linear_solver.jl:155-159) */
            cr[0] = (int64_t)row;
            if (ix > 0) cr[1] = (int64_t)row - 1;
            if (ix < nx - 1) cr[2] = (int64_t)row + 1;
            if (iy > 0) cr[3] = (int64_t)row - nx;
            if (iy < ny - 1) cr[4] = (int64_t)row + nx;
            double m = s->mask[row];
            if (m == OCEAN || m == LAND) { vr[0] = 1.0; rhs[row] = dirichlet_head(m, s->zb[row], p); continue; }
            if (m == OTHER_BASIN) { vr[0] = 1.0; rhs[row] = s->h[row]; continue; }
            double aE = ix < nx - 1 ? boundary_K_face(kfs, s->K, s->mask, row, row + 1) / g->dx2 : 0.0;
            double aW = ix > 0 ? boundary_K_face(kfs, s->K, s->mask, row, row - 1) / g->dx2 : 0.0;
            double aN = iy < ny - 1 ? boundary_K_face(kfs, s->K, s->mask, row, row + nx) / g->dy2 : 0.0;
            double aS = iy > 0 ? boundary_K_face(kfs, s->K, s->mask, row, row - nx) / g->dy2 : 0.0;
            double Np = jl_pow(fabs(s->N[row]), e);
            double aP = parabolic ? (aE + aW + aN + aS) + p->e_v / dt
                                  : (aE + aW + aN + aS) + p->n * p->rho_w * p->g * s->A_visc[row] * Np * s->b[row];
            vr[0] += aP;
            double drhs = 0.0;
            if (ix < nx - 1) { double mo = s->mask[row + 1];
                if (mo == OCEAN || mo == LAND) drhs += aE * dirichlet_head(mo, s->zb[row + 1], p); else vr[2] -= aE; }
            if (ix > 0) { double mo = s->mask[row - 1];
                if (mo == OCEAN || mo == LAND) drhs += aW * dirichlet_head(mo, s->zb[row - 1], p); else vr[1] -= aW; }
            if (iy < ny - 1) { double mo = s->mask[row + nx];
                if (mo == OCEAN || mo == LAND) drhs += aN * dirichlet_head(mo, s->zb[row + nx], p); else vr[4] -= aN; }
            if (iy > 0) { double mo = s->mask[row - nx];
                if (mo == OCEAN || mo == LAND) drhs += aS * dirichlet_head(mo, s->zb[row - nx], p); else vr[3] -= aS; }
            if (!parabolic)
                rhs[row] = s->mdot[row] * (1.0 / p->rho_w - 1.0 / p->rho_i) - s->beta[row] * s->abs_ub[row] +
                           s->A_visc[row] * Np * s->N[row] * s->b[row] +
                           s->A_visc[row] * Np * (p->n * p->rho_w * p->g * s->h[row]) * s->b[row] + s->ieb[row] + drhs;
            else
                rhs[row] = s->mdot[row] * (1.0 / p->rho_w - 1.0 / p->rho_i) - s->beta[row] * s->abs_ub[row] +
                           s->A_visc[row] * Np * s->N[row] * s->b[row] + s->ieb[row] + (p->e_v / dt) * s->h[row] + drhs;
        }
}

/* ---------------------------------------------------------- vector kernels */

/* src/linear_solver.jl:616-669 stencil_matvec_kernel! */
void orc_matvec(int32_t nx, int32_t ny, const double* aP, const double* aE, const double* aW, const double* aN,
                const double* aS, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t c = C_(ix, iy);
            double v = aP[c] * x[c];
            if (ix < nx - 1) v -= aE[c] * x[c + 1];
            if (ix > 0) v -= aW[c] * x[c - 1];
            if (iy < ny - 1) v -= aN[c] * x[c + nx];
            if (iy > 0) v -= aS[c] * x[c - nx];
            y[c] = v;
        }
}

#define NCHUNK 256
static double dot(size_t n, const double* a, const double* b) {
    double part[NCHUNK];
    size_t per = (n + NCHUNK - 1) / NCHUNK;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < NCHUNK; ++k) {
        size_t lo = (size_t)k * per, hi = lo + per < n ? lo + per : n;
        double acc = 0.0;
        for (size_t i = lo; i < hi; ++i) acc += a[i] * b[i];
        part[k] = acc;
    }
    double s = 0.0;
    for (int k = 0; k < NCHUNK; ++k) s += part[k];
    return s;
}
double orc_dot(size_t n, const double* a, const double* b) { return dot(n, a, b); }

/* ----------------------------------------------------- Chebyshev & Lanczos */

typedef struct {
    int32_t nx, ny;
    const double *aP, *aE, *aW, *aN, *aS; /* operator coefficients */
    const double* d;                      /* Jacobi diagonal */
} orc_op;

/* src/preconditioner.jl:66-70 JacobiScaledOperator: y = (A p) ./ d */
static void jac_mul(const orc_op* op, const double* p, double* y) {
    size_t n = (size_t)op->nx * op->ny;
    orc_matvec(op->nx, op->ny, op->aP, op->aE, op->aW, op->aN, op->aS, p, y);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) y[i] /= op->d[i];
}

/* Eigenvalues of a symmetric tridiagonal matrix by Sturm-sequence bisection
 * (textbook; Wilkinson 1965, Golub & Van Loan sec. 8.4.1).  Stands in for
 * LAPACK's eigvals(SymTridiagonal) used at src/preconditioner.jl:125; only the
 * extreme eigenvalues are needed.  Returns 0, or -1 if the input is not finite. */
static int sturm_count(int m, const double* a, const double* e2, double x) {
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < m; ++i) {
        if (q == 0.0) q = 1e-300;
        q = a[i] - x - e2[i - 1] / q;
        if (q < 0) cnt++;
    }
    return cnt;
}
int orc_tridiag_extreme_eigs(int32_t m, const double* a, const double* e, double* lmin, double* lmax) {
    double e2[256];
    if (m < 1 || m > 256) return -1;
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < m; ++i) {
        if (!isfinite(a[i])) { *lmin = *lmax = NAN; return -1; }
        double r = 0.0;
        if (i > 0) r += fabs(e[i - 1]);
        if (i < m - 1) r += fabs(e[i]);
        if (i < m - 1 && !isfinite(e[i])) { *lmin = *lmax = NAN; return -1; }
        lo = fmin(lo, a[i] - r);
        hi = fmax(hi, a[i] + r);
    }
    for (int i = 0; i < m - 1; ++i) e2[i] = e[i] * e[i];
    double span = fmax(fabs(lo), fabs(hi));
    lo -= 1e-14 * span + 1e-300; hi += 1e-14 * span + 1e-300;
    /* smallest: the x where count(x) goes from 0 to >=1 */
    double a0 = lo, b0 = hi;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (a0 + b0);
        if (mid == a0 || mid == b0) break;
        if (sturm_count(m, a, e2, mid) >= 1) b0 = mid; else a0 = mid;
    }
    *lmin = 0.5 * (a0 + b0);
    a0 = lo; b0 = hi;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (a0 + b0);
        if (mid == a0 || mid == b0) break;
        if (sturm_count(m, a, e2, mid) >= m) b0 = mid; else a0 = mid;
    }
    *lmax = 0.5 * (a0 + b0);
    return 0;
}

/* src/preconditioner.jl:89-134 estimate_eigenvalue_bounds: nsteps plain CG
 * steps on D^-1 A with Euclidean dots from x0 = 0, Lanczos tridiagonal from the
 * CG coefficients, Ritz extremes scaled by 0.9 / 1.1. */
void orc_estimate_eigenvalue_bounds(const orc_op* op, const double* rhs, int32_t nsteps, double* lmin, double* lmax) {
    if (nsteps < 1 || nsteps > 256) { *lmin = NAN; *lmax = NAN; return; } /* the Lanczos arrays hold 256 */
    size_t n = (size_t)op->nx * op->ny;
    double* r = malloc(n * sizeof(double));
    double* p = malloc(n * sizeof(double));
    double* Ap = malloc(n * sizeof(double));
    double alphas[256] = {0}, betas[256] = {0};
    memcpy(r, rhs, n * sizeof(double));
    memcpy(p, r, n * sizeof(double));
    double gamma = dot(n, r, r);
    for (int k = 0; k < nsteps; ++k) {
        jac_mul(op, p, Ap);
        double pAp = dot(n, p, Ap);
        alphas[k] = gamma / pAp;
        if (k < nsteps - 1) {
            for (size_t i = 0; i < n; ++i) r[i] -= alphas[k] * Ap[i];
            double gn = dot(n, r, r);
            betas[k] = gn / gamma;
            for (size_t i = 0; i < n; ++i) p[i] = r[i] + betas[k] * p[i];
            gamma = gn;
        }
    }
    double dd[256], ee[256];
    dd[0] = 1.0 / alphas[0];
    for (int k = 1; k < nsteps; ++k) dd[k] = 1.0 / alphas[k] + betas[k - 1] / alphas[k - 1];
    for (int k = 0; k < nsteps - 1; ++k) ee[k] = sqrt(betas[k]) / alphas[k];
    double mn, mx;
    if (orc_tridiag_extreme_eigs(nsteps, dd, ee, &mn, &mx) != 0) { mn = NAN; mx = NAN; }
    *lmin = mn * 0.9;
    *lmax = mx * 1.1;
    free(r); free(p); free(Ap);
}

/* The same estimate in the D-inner product <x, y>_D = x'Dy, where D^-1 A is
 * self-adjoint: CG on D^-1 A x = D^-1 rhs from x0 = 0 (r = rhs ./ d, gamma =
 * r'Dr, p'Ap -> p'(A p)), i.e. the coefficients of Jacobi-preconditioned CG,
 * a symmetric Lanczos process whose Ritz values lie in the spectrum of
 * D^-1 A.  Scaled by 0.9 / 1.1 as the reference scales its own.  The HIP
 * library's Gershgorin bound mode takes lambda_min from it. */
void orc_estimate_eigenvalue_bounds_d(const orc_op* op, const double* rhs, int32_t nsteps, double* lmin,
                                      double* lmax) {
    if (nsteps < 1 || nsteps > 256) { *lmin = NAN; *lmax = NAN; return; } /* the Lanczos arrays hold 256 */
    size_t n = (size_t)op->nx * op->ny;
    double* r = malloc(n * sizeof(double));
    double* p = malloc(n * sizeof(double));
    double* Ap = malloc(n * sizeof(double));
    double* Au = malloc(n * sizeof(double));
    double alphas[256] = {0}, betas[256] = {0};
    for (size_t i = 0; i < n; ++i) r[i] = rhs[i] / op->d[i];
    memcpy(p, r, n * sizeof(double));
    double gamma = 0.0;
    for (size_t i = 0; i < n; ++i) gamma += r[i] * r[i] * op->d[i];
    for (int k = 0; k < nsteps; ++k) {
        orc_matvec(op->nx, op->ny, op->aP, op->aE, op->aW, op->aN, op->aS, p, Au);
        double pAp = dot(n, p, Au);
        for (size_t i = 0; i < n; ++i) Ap[i] = Au[i] / op->d[i];
        alphas[k] = gamma / pAp;
        if (k < nsteps - 1) {
            double gn = 0.0;
            for (size_t i = 0; i < n; ++i) {
                r[i] -= alphas[k] * Ap[i];
                gn += r[i] * r[i] * op->d[i];
            }
            betas[k] = gn / gamma;
            for (size_t i = 0; i < n; ++i) p[i] = r[i] + betas[k] * p[i];
            gamma = gn;
        }
    }
    double dd[256], ee[256];
    dd[0] = 1.0 / alphas[0];
    for (int k = 1; k < nsteps; ++k) dd[k] = 1.0 / alphas[k] + betas[k - 1] / alphas[k - 1];
    for (int k = 0; k < nsteps - 1; ++k) ee[k] = sqrt(betas[k]) / alphas[k];
    double mn, mx;
    if (orc_tridiag_extreme_eigs(nsteps, dd, ee, &mn, &mx) != 0) { mn = NAN; mx = NAN; }
    *lmin = mn * 0.9;
    *lmax = mx * 1.1;
    free(r); free(p); free(Ap); free(Au);
}

/* src/preconditioner.jl:197-233 update_chebyshev_bounds!: adopt the estimate
 * only if finite, positive and non-empty; otherwise keep the previous bounds
 * and report a warning (return 1). */
int orc_update_chebyshev_bounds(const orc_op* op, const double* rhs, int32_t nsteps, double* lmin, double* lmax) {
    double a, b;
    orc_estimate_eigenvalue_bounds(op, rhs, nsteps, &a, &b);
    if (isfinite(a) && isfinite(b) && a > 0 && b > a) { *lmin = a; *lmax = b; return 0; }
    return 1;
}

/* src/preconditioner.jl:239-267 ldiv!(y, ::ChebyshevPreconditioner, x) (Saad Alg. 12.1) */
void orc_cheb_apply(const orc_op* op, int32_t degree, double lmin, double lmax, const double* x, double* y,
                    double* r, double* p, double* Ap) {
    size_t n = (size_t)op->nx * op->ny;
    double d_mid = (lmax + lmin) / 2.0, c_rad = (lmax - lmin) / 2.0;
    for (size_t i = 0; i < n; ++i) { r[i] = x[i] / op->d[i]; y[i] = 0.0; }
    double alpha = 1.0 / d_mid;
    for (size_t i = 0; i < n; ++i) { p[i] = r[i]; y[i] += alpha * p[i]; }
    jac_mul(op, p, Ap);
    for (size_t i = 0; i < n; ++i) r[i] -= alpha * Ap[i];
    double alpha_prev = alpha;
    for (int k = 2; k <= degree; ++k) {
        double t = c_rad * alpha_prev / 2.0;
        double beta = t * t;
        alpha = 1.0 / (d_mid - beta / alpha_prev);
        for (size_t i = 0; i < n; ++i) { p[i] = r[i] + beta * p[i]; y[i] += alpha * p[i]; }
        jac_mul(op, p, Ap);
        for (size_t i = 0; i < n; ++i) r[i] -= alpha * Ap[i];
        alpha_prev = alpha;
    }
}

/* Gershgorin bound of D^-1 A over the couplings the operator applies:
 * 1 + max_i sum_j |a_ij| / |d_i| (every eigenvalue of D^-1 A lies in
 * [1 - rho, 1 + rho], rho = that max - 1; Gershgorin's circle theorem).  NaN
 * when a row has a zero diagonal and no couplings. */
double orc_gershgorin(const orc_op* op) {
    const int32_t nx = op->nx, ny = op->ny;
    double m = -INFINITY;
    int nan = 0;
    for (int32_t iy = 0; iy < ny; ++iy)
        for (int32_t ix = 0; ix < nx; ++ix) {
            const size_t c = (size_t)ix + (size_t)iy * nx;
            double s = 0.0;
            if (ix < nx - 1) s += fabs(op->aE[c]);
            if (ix > 0) s += fabs(op->aW[c]);
            if (iy < ny - 1) s += fabs(op->aN[c]);
            if (iy > 0) s += fabs(op->aS[c]);
            const double v = 1.0 + s / fabs(op->d[c]);
            if (isnan(v)) nan = 1;
            else if (v > m) m = v;
        }
    return nan ? NAN : m;
}

/* The bounds of one solve in the Gershgorin mode: lambda_min from the
 * D-inner-product estimate, lambda_max = G, adopted when finite with
 * 0 < lambda_min < G; otherwise the previous lambda_min with this G (the
 * placeholder (1, 1) stays as it is).  Returns 1 when the estimate was
 * rejected. */
int orc_gershgorin_bounds(const orc_op* op, const double* rhs, int32_t nsteps, double* lmin, double* lmax) {
    double a, b;
    orc_estimate_eigenvalue_bounds_d(op, rhs, nsteps, &a, &b);
    const double G = orc_gershgorin(op);
    const int gok = isfinite(G) && G > 0;
    if (gok && isfinite(a) && a > 0 && G > a) { *lmin = a; *lmax = G; return 0; }
    if (gok && *lmin != *lmax && G > *lmin) *lmax = G;
    return 1;
}

/* ------------------------------------------------------------ CG solver */

typedef struct {
    int32_t precond;       /* 0 Jacobi, 1 Chebyshev */
    int32_t cheb_degree;
    int32_t cheb_nsteps;
    int32_t itmax;         /* 0 -> 2n, Krylov's default */
    double atol, rtol;     /* Krylov defaults are sqrt(eps) */
    double lambda_min, lambda_max; /* persistent Chebyshev bounds; start at (1, 1) */
    int32_t gershgorin;    /* 0: the reference's bounds; 1: lambda_max from the Gershgorin bound (the HIP
                              library's default, a stated deviation from the reference) */
    int32_t pad;
} orc_cg_cfg;

typedef struct {
    int32_t niter, solved, zero_curvature, bounds_warning;
    double rnorm0, rnorm, lambda_min, lambda_max;
    int32_t breakdown; /* 1: r'z < 0 or not finite, where Krylov's sqrt(r'z) throws (x0 left unchanged) */
    int32_t pad;
} orc_cg_stats;

/* Krylov.jl 0.10.8 cg! (not vendored; restated from its published algorithm)
 * as called at src/linear_solver.jl:915-945: warm start x0 (the current h),
 * preconditioner applied with ldiv, residual measured in the preconditioner
 * norm sqrt(r'z), stop when rNorm <= atol + rtol*rNorm0 or rNorm + 1 <= 1,
 * zero curvature when |p'Ap| <= eps*||p||^2_M, itmax = 2n by default.
 * x0 is overwritten with x0 + dx (the reference's s.h .= ws.x).  A negative
 * or non-finite r'z stops the solve with st->breakdown = 1, x0 untouched and
 * a return value of -2: Julia's sqrt throws a DomainError there, so the
 * reference's cg! never returns and s.h keeps its value. */
int orc_cg_solve(const orc_op* op, orc_cg_cfg* cfg, const double* b, double* x0, orc_cg_stats* st) {
    size_t n = (size_t)op->nx * op->ny;
    double *x = calloc(n, sizeof(double)), *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double));
    double *p = malloc(n * sizeof(double)), *Ap = malloc(n * sizeof(double));
    double *cr = malloc(n * sizeof(double)), *cp = malloc(n * sizeof(double)), *cAp = malloc(n * sizeof(double));
    memset(st, 0, sizeof(*st));
    if (cfg->precond == 1)
        st->bounds_warning =
            cfg->gershgorin
                ? orc_gershgorin_bounds(op, b, cfg->cheb_nsteps, &cfg->lambda_min, &cfg->lambda_max)
                : orc_update_chebyshev_bounds(op, b, cfg->cheb_nsteps, &cfg->lambda_min, &cfg->lambda_max);
    st->lambda_min = cfg->lambda_min;
    st->lambda_max = cfg->lambda_max;
#define PRECOND(zz, rr)                                                                                      \
    do {                                                                                                     \
        if (cfg->precond == 1)                                                                               \
            orc_cheb_apply(op, cfg->cheb_degree, cfg->lambda_min, cfg->lambda_max, rr, zz, cr, cp, cAp);     \
        else                                                                                                 \
            for (size_t i_ = 0; i_ < n; ++i_) zz[i_] = rr[i_] / op->d[i_];                                  \
    } while (0)
    /* warm start: r = b - A x0 */
    orc_matvec(op->nx, op->ny, op->aP, op->aE, op->aW, op->aN, op->aS, x0, r);
    for (size_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    PRECOND(z, r);
    double gamma = dot(n, r, z);
    double rNorm = sqrt(gamma);
    st->rnorm0 = rNorm;
    if (!(gamma >= 0.0) || !isfinite(gamma)) {
        st->breakdown = 1;
        st->rnorm = rNorm;
        goto fail;
    }
    if (gamma == 0.0) {
        st->solved = 1; st->rnorm = rNorm; st->niter = 0;
        goto done;
    }
    {
        long itmax = cfg->itmax > 0 ? cfg->itmax : 2 * (long)n;
        double pNorm2 = gamma;
        double eps_stop = cfg->atol + cfg->rtol * rNorm;
        memcpy(p, z, n * sizeof(double));
        int solved = rNorm <= eps_stop, tired = 0, zc = 0;
        long iter = 0;
        while (!(solved || tired || zc)) {
            orc_matvec(op->nx, op->ny, op->aP, op->aE, op->aW, op->aN, op->aS, p, Ap);
            double pAp = dot(n, p, Ap);
            if (pAp <= EPS64 * pNorm2 && fabs(pAp) <= EPS64 * pNorm2) { zc = 1; break; }
            double alpha = gamma / pAp;
            for (size_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
            PRECOND(z, r);
            double gn = dot(n, r, z);
            rNorm = sqrt(gn);
            if (!(gn >= 0.0) || !isfinite(gn)) {
                st->niter = (int32_t)(iter + 1);
                st->breakdown = 1;
                st->rnorm = rNorm;
                goto fail;
            }
            solved = (rNorm <= eps_stop) || (rNorm + 1.0 <= 1.0);
            if (!solved) {
                double beta = gn / gamma;
                pNorm2 = gn + beta * beta * pNorm2;
                gamma = gn;
                for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            }
            iter++;
            tired = iter >= itmax;
        }
        st->niter = (int32_t)iter;
        st->solved = solved;
        st->zero_curvature = zc;
        st->rnorm = rNorm;
    }
done:
    for (size_t i = 0; i < n; ++i) x0[i] = x[i] + x0[i];
fail:
#undef PRECOND
    free(x); free(r); free(z); free(p); free(Ap); free(cr); free(cp); free(cAp);
    return st->breakdown ? -2 : 0;
}

/* --------------------------------------------- banded Cholesky (oracle) */

/* Dense-band Cholesky of the assembled SPD operator: the CHOLMOD
 * cholesky!/ldiv! of src/linear_solver.jl:709-739, restated as LAPACK-style
 * band factorisation (dpbtrf/dpbtrs order of operations).  The unknowns are
 * renumbered iy-fastest when ny < nx so the half-bandwidth is min(nx, ny).
 * Returns 0, or the 1-based pivot that failed (not SPD). */
int orc_cholesky_solve(int32_t nx, int32_t ny, const double* aP, const double* aE, const double* aW,
                       const double* aN, const double* aS, const double* rhs, double* x) {
    const size_t n = (size_t)nx * ny;
    const int swap = ny < nx;
    const int bw = swap ? ny : nx;
    double* L = calloc(n * (size_t)(bw + 1), sizeof(double)); /* L[j*(bw+1) + (i-j)] = L(i,j), i in [j, j+bw] */
    double* y = malloc(n * sizeof(double));
#define PERM(ix, iy) (swap ? ((size_t)(iy) + (size_t)(ix) * ny) : ((size_t)(ix) + (size_t)(iy) * nx))
#define LB(i, j) L[(size_t)(j) * (bw + 1) + ((i) - (j))]
    /* scatter lower triangle: for new row index r, entries at columns c <= r */
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            size_t c = C_(ix, iy), r = PERM(ix, iy);
            LB(r, r) = aP[c];
            /* coupling (row r, col o) with o < r: entry -a */
            if (ix > 0) { size_t o = PERM(ix - 1, iy); if (o < r) LB(r, o) = -aW[c]; }
            if (ix < nx - 1) { size_t o = PERM(ix + 1, iy); if (o < r) LB(r, o) = -aE[c]; }
            if (iy > 0) { size_t o = PERM(ix, iy - 1); if (o < r) LB(r, o) = -aS[c]; }
            if (iy < ny - 1) { size_t o = PERM(ix, iy + 1); if (o < r) LB(r, o) = -aN[c]; }
        }
    int info = 0;
    for (size_t j = 0; j < n && !info; ++j) {
        size_t k0 = j >= (size_t)bw ? j - bw : 0;
        double s = LB(j, j);
        for (size_t k = k0; k < j; ++k) s -= LB(j, k) * LB(j, k);
        if (!(s > 0.0)) { info = (int)j + 1; break; }
        double ljj = sqrt(s);
        LB(j, j) = ljj;
        size_t imax = j + bw < n - 1 ? j + bw : n - 1;
#pragma omp parallel for schedule(static) if (bw > 64)
        for (size_t i = j + 1; i <= imax; ++i) {
            size_t kk0 = i >= (size_t)bw ? i - bw : 0;
            if (kk0 < k0) kk0 = k0;
            double v = LB(i, j);
            for (size_t k = kk0; k < j; ++k) v -= LB(i, k) * LB(j, k);
            LB(i, j) = v / ljj;
        }
    }
    if (!info) {
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) y[PERM(ix, iy)] = rhs[C_(ix, iy)];
        for (size_t i = 0; i < n; ++i) {
            size_t k0 = i >= (size_t)bw ? i - bw : 0;
            double v = y[i];
            for (size_t k = k0; k < i; ++k) v -= LB(i, k) * y[k];
            y[i] = v / LB(i, i);
        }
        for (size_t ii = n; ii-- > 0;) {
            size_t imax = ii + bw < n - 1 ? ii + bw : n - 1;
            double v = y[ii];
            for (size_t k = ii + 1; k <= imax; ++k) v -= LB(k, ii) * y[k];
            y[ii] = v / LB(ii, ii);
        }
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) x[C_(ix, iy)] = y[PERM(ix, iy)];
    }
#undef PERM
#undef LB
    free(L); free(y);
    return info;
}

/* -------------------------------------------------- solver front-ends */

typedef struct {
    int32_t kind;          /* 0 Cholesky (oracle), 1 matrix-free CG */
    orc_cg_cfg cg;
    double *aP, *aE, *aW, *aN, *aS, *rhs; /* owned workspace */
    int32_t nx, ny;
} orc_solver;

orc_solver* orc_solver_create(int32_t nx, int32_t ny, int32_t kind, const orc_cg_cfg* cg) {
    orc_solver* s = calloc(1, sizeof(orc_solver));
    size_t n = (size_t)nx * ny;
    s->kind = kind; s->nx = nx; s->ny = ny;
    if (cg) s->cg = *cg;
    if (s->cg.lambda_min == 0.0 && s->cg.lambda_max == 0.0) { s->cg.lambda_min = 1.0; s->cg.lambda_max = 1.0; }
    s->aP = malloc(n * sizeof(double)); s->aE = malloc(n * sizeof(double)); s->aW = malloc(n * sizeof(double));
    s->aN = malloc(n * sizeof(double)); s->aS = malloc(n * sizeof(double)); s->rhs = malloc(n * sizeof(double));
    return s;
}
void orc_solver_destroy(orc_solver* s) {
    if (!s) return;
    free(s->aP); free(s->aE); free(s->aW); free(s->aN); free(s->aS); free(s->rhs); free(s);
}
void orc_solver_bounds(const orc_solver* s, double* lmin, double* lmax) { *lmin = s->cg.lambda_min; *lmax = s->cg.lambda_max; }

/* solve_elliptic_linear_system! / solve_parabolic_linear_system!
 * (src/linear_solver.jl:709-739 Cholesky, :915-945 matrix-free CG) */
int orc_solve_linear(orc_solver* sv, const orc_grid* g, const orc_params* p, int32_t kfs, orc_state* s,
                     int32_t parabolic, double dt, orc_cg_stats* st) {
    orc_assemble_mfls(g, p, kfs, s, parabolic, dt, sv->aP, sv->aE, sv->aW, sv->aN, sv->aS, sv->rhs);
    if (sv->kind == 0) {
        if (st) memset(st, 0, sizeof(*st));
        int info = orc_cholesky_solve(g->nx, g->ny, sv->aP, sv->aE, sv->aW, sv->aN, sv->aS, sv->rhs, s->h);
        if (st) st->solved = info == 0;
        return info == 0 ? 0 : -1;
    }
    orc_op op = {g->nx, g->ny, sv->aP, sv->aE, sv->aW, sv->aN, sv->aS, sv->aP};
    orc_cg_stats tmp;
    return orc_cg_solve(&op, &sv->cg, sv->rhs, s->h, st ? st : &tmp);
}

/* src/elliptic_solver.jl:138-208 Picard_loop!/Picard_iteration!/relax_h!.
 * relax_alpha <= 0 means NoHeadRelaxation.  Returns converged (-1 when a
 * linear solve failed); *last_iter as the reference's ps.last_iter.
 * *cg_iters accumulates CG iterations. */
int32_t orc_picard_loop(orc_solver* sv, const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s,
                        int32_t iters, double tol, int32_t check_every, double relax_alpha, int32_t* last_iter,
                        int64_t* cg_iters) {
    size_t n = (size_t)g->nx * g->ny;
    double* h_prev = malloc(n * sizeof(double));
    int32_t converged = 0;
    *last_iter = 0;
    for (int32_t it = 1; it <= iters; ++it) {
        memcpy(h_prev, s->h, n * sizeof(double));
        orc_cg_stats st;
        if (orc_solve_linear(sv, g, p, m->kfs, s, 0, 0.0, &st) != 0) {
            /* the reference's solve throws (CG breakdown, failed factorisation) */
            free(h_prev);
            return -1;
        }
        if (cg_iters) *cg_iters += st.niter;
        if (relax_alpha > 0.0)
            for (size_t i = 0; i < n; ++i) s->h[i] = relax_alpha * s->h[i] + (1.0 - relax_alpha) * h_prev[i];
        orc_update_derived(g, p, m, s);
        if (it % check_every == 0 || it == iters) {
            /* mapreduce(abs, max, ...; init = 0): any NaN makes the maximum NaN */
            double dmax = 0.0, hmax = 0.0;
            int any_nan = 0;
            for (size_t i = 0; i < n; ++i) {
                double d = fabs(s->h[i] - h_prev[i]), a = fabs(s->h[i]);
                if (isnan(d) || isnan(a)) any_nan = 1;
                if (d > dmax) dmax = d;
                if (a > hmax) hmax = a;
            }
            if (!any_nan && dmax / (hmax + EPS64) < tol) { converged = 1; *last_iter = it; break; }
        }
    }
    if (!converged) *last_iter = iters;
    free(h_prev);
    return converged;
}

/* src/parabolic_solver.jl:16-38 */
int orc_parabolic_step(orc_solver* sv, const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s,
                       double dt, int64_t* cg_iters) {
    orc_cg_stats st;
    int rc = orc_solve_linear(sv, g, p, m->kfs, s, 1, dt, &st);
    if (cg_iters) *cg_iters += st.niter;
    orc_update_derived(g, p, m, s);
    return rc;
}

/* src/melt_input.jl:86-94 update_ieb!(::SeasonalMeltInput): the uniform value
 * for elapsed time t (seconds).  Returned, not applied. */
double orc_seasonal_ieb(double t, double t_start, double period, double amplitude, double offset, double i_min,
                        double seconds_per_year) {
    double omega = 2.0 * M_PI / period;
    double yf = fmod(t / seconds_per_year, 1.0);
    if (yf < 0) yf += 1.0;
    double i_ma = (t_start <= yf && yf <= t_start + period) ? (offset - amplitude * cos(omega * (yf - t_start))) : i_min;
    return i_ma / seconds_per_year;
}

/* One whole timestep, src/run.jl:82-88 (step!): ieb refresh (mode 0 constant:
 * untouched; mode 1: uniform overwrite; mode 2: uniform + static moulin field),
 * step_h! (elliptic Picard or parabolic), step_b!.  Returns converged (or 1
 * for parabolic), writes Picard's last_iter. */
int32_t orc_step(orc_solver* sv, const orc_grid* g, const orc_params* p, const orc_model* m, orc_state* s, double dt,
                 int32_t ieb_mode, double ieb_uniform, const double* moulin, int32_t iters, double tol,
                 int32_t check_every, double relax_alpha, int32_t* last_iter, int64_t* cg_iters) {
    size_t n = (size_t)g->nx * g->ny;
    if (ieb_mode == 1) for (size_t i = 0; i < n; ++i) s->ieb[i] = ieb_uniform;
    else if (ieb_mode == 2) for (size_t i = 0; i < n; ++i) s->ieb[i] = ieb_uniform + moulin[i];
    int32_t conv = 1;
    if (p->e_v == 0.0) {
        conv = orc_picard_loop(sv, g, p, m, s, iters, tol, check_every, relax_alpha, last_iter, cg_iters);
        if (conv < 0) return -1;
    } else {
        if (orc_parabolic_step(sv, g, p, m, s, dt, cg_iters) != 0) return -1;
        *last_iter = 0;
    }
    orc_step_b(g, p, m, s, dt);
    return conv;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * The synthetic glacier of shk_generate_glacier, restated on the CPU (no
 * reference counterpart: the reference takes these inputs as arguments,
 * src/initial_conditions.jl:14-126).  Same stream, same loops, same order of
 * every sum, so the fields agree bit for bit:
 *   o_i = SplitMix64 output i from state seed (Vigna's generator);
 *   noise w_c = sum_{k<12} (o_{12c+k} >> 16) 2^-48 - 6;
 *   periodic Gaussian smoothing along x then y (weights exp(-(k/s)^2/2)
 *   normalised over k = -R..R, s = corr_len / (sqrt(2) d), R = ceil(4 s));
 *   zero mean and standard deviation bed_amp (sums over 16384 interleaved
 *   chunks, then 128 groups of 128 chunk sums, then the groups);
 *   Gaussian trough; zs = 6 (sqrt(x + 5000) - sqrt(5000)) + 1; uniform fields;
 *   moulins on distinct interior cells.
 * ------------------------------------------------------------------------ */
typedef struct {
    uint64_t seed;
    double bed_amp, corr_len, trough_width, trough_depth;
    double b0, A_visc, ub, G, ieb;
    int32_t n_moulins, reserved;
    double q_min, q_max;
} orc_glacier_spec;

#define GEN_PHI 0x9E3779B97F4A7C15ull
#define GEN_NCHUNK 16384

static uint64_t gen_splitmix64(uint64_t z) {
    z += GEN_PHI;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t z) { return gen_splitmix64(z); }

static double gen_unit53(uint64_t o) { return (double)(o >> 11) * 0x1p-53; }

/* exp(-(k/s)^2/2) normalised over k = -R..R; returns R (weights in *w, malloc'd) */
static int gen_weights(double s, double** w) {
    if (!(s > 0.0)) {
        *w = malloc(sizeof(double));
        (*w)[0] = 1.0;
        return 0;
    }
    const int R = (int)ceil(4.0 * s);
    double* v = malloc((size_t)(2 * R + 1) * sizeof(double));
    double sum = 0.0;
    for (int k = -R; k <= R; ++k) {
        const double t = (double)k / s;
        v[k + R] = exp(-0.5 * t * t);
        sum = sum + v[k + R];
    }
    for (int k = 0; k <= 2 * R; ++k) v[k] = v[k] / sum;
    *w = v;
    return R;
}

static double gen_sum(const double* f, size_t n, double mean, int square) {
    double* part = malloc(GEN_NCHUNK * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < GEN_NCHUNK; ++c) {
        double acc = 0.0;
        for (size_t i = (size_t)c; i < n; i += GEN_NCHUNK) {
            double v = f[i];
            if (square) {
                v = v - mean;
                v = v * v;
            }
            acc = acc + v;
        }
        part[c] = acc;
    }
    double grp[128];
    for (int t = 0; t < 128; ++t) {
        double acc = 0.0;
        for (int c = 0; c < 128; ++c) acc = acc + part[128 * t + c];
        grp[t] = acc;
    }
    double s = 0.0;
    for (int t = 0; t < 128; ++t) s = s + grp[t];
    free(part);
    return s;
}

/* Writes mask, A_visc, zb, zs, b, G, ub_x, ub_y, ieb, taub_x, taub_y of s and,
 * when moulin is not NULL, the moulin field.  Returns 0, or -1 on bad input. */
int orc_generate_glacier(const orc_grid* g, const orc_glacier_spec* sp, orc_state* s, double* moulin) {
    const int nx = g->nx, ny = g->ny;
    const size_t n = (size_t)nx * ny;
    if (nx < 2 || ny < 2) return -1;
    if (sp->n_moulins > 0 && (nx < 3 || ny < 3 || (int64_t)sp->n_moulins > (int64_t)(nx - 2) * (ny - 2))) return -1;
    double* w = malloc(n * sizeof(double));
    double* tmp = malloc(n * sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        const uint64_t base = sp->seed + (uint64_t)(12 * c) * GEN_PHI;
        double acc = 0.0;
        for (int k = 0; k < 12; ++k) acc = acc + (double)(gen_splitmix64(base + (uint64_t)k * GEN_PHI) >> 16) * 0x1p-48;
        w[c] = acc - 6.0;
    }
    double *wx, *wy;
    const int Rx = gen_weights(sp->corr_len / (sqrt(2.0) * g->dx), &wx);
    const int Ry = gen_weights(sp->corr_len / (sqrt(2.0) * g->dy), &wy);
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            int j = (ix - Rx) % nx;
            if (j < 0) j += nx;
            double acc = 0.0;
            for (int k = 0; k <= 2 * Rx; ++k) {
                acc = acc + wx[k] * w[(size_t)iy * nx + j];
                if (++j == nx) j = 0;
            }
            tmp[(size_t)iy * nx + ix] = acc;
        }
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            int j = (iy - Ry) % ny;
            if (j < 0) j += ny;
            double acc = 0.0;
            for (int k = 0; k <= 2 * Ry; ++k) {
                acc = acc + wy[k] * tmp[(size_t)j * nx + ix];
                if (++j == ny) j = 0;
            }
            w[(size_t)iy * nx + ix] = acc;
        }
    const double mean = gen_sum(w, n, 0.0, 0) / (double)n;
    const double sd = sqrt(gen_sum(w, n, mean, 1) / (double)n);
    const double scale = sd > 0.0 ? sp->bed_amp / sd : 1.0;
    const int has_trough = sp->trough_width > 0.0 && sp->trough_depth != 0.0;
    uint64_t next = 12 * (uint64_t)n;
    double y0 = 0.0;
    if (has_trough) y0 = (0.25 + 0.5 * gen_unit53(gen_splitmix64(sp->seed + next * GEN_PHI))) * g->ly;
    ++next;
    const double sig = sp->trough_width / 2.0;
    for (int iy = 0; iy < ny; ++iy) {
        double prof = 0.0;
        if (has_trough) {
            const double t = ((double)iy * g->dy - y0) / sig;
            prof = exp(-0.5 * t * t);
        }
        for (int ix = 0; ix < nx; ++ix) {
            const size_t c = C_(ix, iy);
            const double d = w[c] - mean;
            double z = sd > 0.0 ? d * scale : d;
            if (has_trough) z = z - sp->trough_depth * prof;
            s->zb[c] = z;
            s->zs[c] = 6.0 * (sqrt((double)ix * g->dx + 5000.0) - sqrt(5000.0)) + 1.0;
            s->mask[c] = ix == 0 ? OCEAN : GROUNDED;
            s->A_visc[c] = sp->A_visc;
            s->b[c] = sp->b0;
            s->G[c] = sp->G;
            s->ieb[c] = sp->ieb;
        }
    }
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
            s->ub_x[XF_(ix, iy)] = sp->ub;
            s->taub_x[XF_(ix, iy)] = 0.0;
        }
    for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            s->ub_y[YF_(ix, iy)] = 0.0;
            s->taub_y[YF_(ix, iy)] = 0.0;
        }
    if (moulin && sp->n_moulins > 0) {
        memset(moulin, 0, n * sizeof(double));
        const int64_t mx = nx - 2, M = (int64_t)(nx - 2) * (ny - 2);
        int64_t* cells = malloc((size_t)sp->n_moulins * sizeof(int64_t));
        for (int k = 0; k < sp->n_moulins; ++k) {
            int64_t c;
            for (;;) {
                const int64_t j = (int64_t)(gen_unit53(gen_splitmix64(sp->seed + next++ * GEN_PHI)) * (double)M);
                c = (1 + j % mx) + (1 + j / mx) * (int64_t)nx;
                int taken = 0;
                for (int q = 0; q < k; ++q) taken |= cells[q] == c;
                if (!taken) break;
            }
            const double u = gen_unit53(gen_splitmix64(sp->seed + next++ * GEN_PHI));
            cells[k] = c;
            moulin[c] = (sp->q_min + (sp->q_max - sp->q_min) * u) / (g->dx * g->dy);
        }
        free(cells);
    }
    free(w);
    free(tmp);
    free(wx);
    free(wy);
    return 0;
}
