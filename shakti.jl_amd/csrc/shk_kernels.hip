// gfx950 kernels of libshakti_hip.so.
//
// Memory layout: every field is column-major with ix fastest (reference
// src/state.jl:13-57): centre (nx, ny) at ix + iy*nx, x-faces (nx+1, ny) at
// ix + iy*(nx+1), y-faces (nx, ny+1) at ix + iy*nx.
//
// The linear system is stored in the symmetric face layout: aP (diagonal),
// aE (coupling of cell (ix,iy) to (ix+1,iy)) and aN (coupling to (ix,iy+1)).
// The reference's aW(ix,iy) equals aE(ix-1,iy) bit for bit and aS(ix,iy)
// equals aN(ix,iy-1) (both faces evaluate the same commutative K_face
// expression, and every coupling to a non-GROUNDED cell is zero on both
// sides), so the operator apply reads 3 coefficient arrays instead of 5.
//
// Stencil kernels march: each wave owns a 64-wide column strip and walks R
// rows upwards, keeping the operand of rows iy-1, iy, iy+1 and aN(iy-1) in
// registers, so every array is read from HBM once per launch (plus one halo
// row per R rows).  Workgroups are renumbered so that the blocks one XCD
// receives (the hardware deals workgroups round-robin over the 8 XCDs) form
// one contiguous band of rows, keeping shared halo rows in that XCD's L2.
//
// Reductions (dot products, maxima) are deterministic: a fixed wave-shuffle
// tree, a fixed cross-wave order, one partial per workgroup, and a
// one-workgroup finishing kernel launched right after on the same stream sums
// the partials in a fixed order and updates the device-resident solver
// scalars.
//
// Floating-point expressions keep the reference's evaluation order and the
// library is built with -ffp-contract=off, so the FP64 path rounds exactly as
// the Julia source is written (only the non-integer powers may differ in the
// last ulps: pow(), and the cbrt form of x^(1/3) in pow_third).

#include <math.h>
#include <stdlib.h>

#include <type_traits>

#include "shk_internal.hpp"

namespace shk {

// --------------------------------------------------------------- helpers

// Julia's Base.:^(::Float64, ::Integer) (power by squaring with an
// error-compensated product, n == 3 as x*x*x); n == 2 reduces to x*x.  The
// body for n >= 0 is not recursive, so the compiler inlines it (a recursive
// form is a real call, with its stack in scratch memory, in every kernel
// that uses it).
__device__ __forceinline__ double jl_pow_nonneg(double x, int n) {
    if (n == 2) return x * x;
    if (n == 3) return x * x * x;
    if (n == 1) return x;
    if (n == 0) return 1.0;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = y * xnlo + x * ynlo;
            double yy = x * y;
            ynlo = fma(x, y, -yy);
            y = yy;
            ynlo += err;
        }
        double err = x * 2.0 * xnlo;
        double xx = x * x;
        xnlo = fma(x, x, -xx);
        x = xx;
        xnlo += err;
        n >>= 1;
    }
    double err = y * xnlo + x * ynlo;
    if (isfinite(x) && isfinite(err)) return fma(x, y, err);
    return x * y;
}
__device__ __forceinline__ double jl_pow_int(double x, int n) {
    return n < 0 ? 1.0 / jl_pow_nonneg(x, -n) : jl_pow_nonneg(x, n);
}

// x^e for e the double nearest 1/3, which is 1/3 - d with d = 2^-54 / 3
// exactly: x^e = cbrt(x) x^-d = cbrt(x) exp(-d ln x), and with |d ln x| below
// 1.4e-14 for every positive double, cbrt(x) (1 - d ln x) to within 1e-28 --
// a correction of at most some 60 ulps, for which the FP32 logarithm of the
// mantissa is more than accurate enough.  A seventh of pow's instructions;
// within 2 ulps of pow (measured over 2^24 arguments per binade range from
// 2^-1074 to 1).  Zero, negative, infinite and NaN arguments give pow's
// values: +0, NaN, +inf, NaN.
__device__ __forceinline__ double pow_third(double x) {
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : NAN;
    if (isinf(x)) return x;
    int ex;
    const double m = frexp(x, &ex);  // x = m 2^ex, m in [0.5, 1)
    const double lnx = ((double)ex + (double)__log2f((float)m)) * 0.69314718055994530942;
    const double y = cbrt(x);
    return y + y * (-lnx * (1.0 / 54043195528445952.0));  // 1 / (3 2^54)
}
__device__ __forceinline__ double cpow(double x, const Expo& e) {
    return e.is_int ? jl_pow_int(x, e.k) : (e.third ? pow_third(x) : pow(x, e.e));
}

__device__ __forceinline__ double jl_sign(double x) { return (x == 0.0 || isnan(x)) ? x : copysign(1.0, x); }
__device__ __forceinline__ double jl_boolmul(bool t, double x) { return t ? x : copysign(0.0, x); }
__device__ __forceinline__ double jl_min(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}
__device__ __forceinline__ double jl_max0(double v) { return isnan(v) ? v : (v > 0.0 ? v : 0.0); }
__device__ __forceinline__ double jl_clamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

__device__ __forceinline__ size_t cidx(int ix, int iy, int nx) { return (size_t)ix + (size_t)iy * (size_t)nx; }
// FP32 -> BF16 bits, round to nearest even
__device__ __forceinline__ uint16_t bf16_rne(float v) {
    const uint32_t u = __float_as_uint(v);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ size_t xidx(int ix, int iy, int nx) { return (size_t)ix + (size_t)iy * (size_t)(nx + 1); }

// Virtual workgroup id: XCD k (hardware id % 8) receives one contiguous range.
__device__ __forceinline__ void virtual_block(int& vbx, int& vby) {
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.x + blockIdx.y * gridDim.x;
    const unsigned q = nb / 8, r = nb % 8, xcd = bid % 8, local = bid / 8;
    const unsigned v = xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
    vbx = (int)(v % gridDim.x);
    vby = (int)(v / gridDim.x);
}

// Wave-wide lane shifts through DPP (GFX9 wave_shr:1 / wave_shl:1): the value
// of lane i-1 (from the west) or lane i+1 (from the east); lanes without a
// source receive 0.
__device__ __forceinline__ int dpp_west_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_east_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ double from_west(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_west_i((int)(b & 0xffffffffll)), hi = dpp_west_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_east(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_east_i((int)(b & 0xffffffffll)), hi = dpp_east_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float from_west(float v) { return __int_as_float(dpp_west_i(__float_as_int(v))); }
__device__ __forceinline__ float from_east(float v) { return __int_as_float(dpp_east_i(__float_as_int(v))); }
// One-lane shifts with the edge semantics of __shfl_up/__shfl_down(v, 1):
// the edge lane without a source keeps its own value (DPP with bound_ctrl
// off writes nothing there and old = v).  A DPP move is one VALU instruction,
// where the shuffle is an LDS permute with its round trip.
__device__ __forceinline__ int lane_up1_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int lane_down1_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ float lane_up1(float v) { return __int_as_float(lane_up1_i(__float_as_int(v))); }
__device__ __forceinline__ float lane_down1(float v) { return __int_as_float(lane_down1_i(__float_as_int(v))); }
__device__ __forceinline__ double lane_up1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_up1_i((int)(b & 0xffffffffll)), hi = lane_up1_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_down1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = lane_down1_i((int)(b & 0xffffffffll)), hi = lane_down1_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, WAVE);
        v = (isnan(o) || o > v) ? o : v;
    }
    return v;
}

// Sum (or max, with MAX) of up to two values over the workgroup; result valid
// in thread 0.  Fixed order, so deterministic.
template <int NV, bool MAX>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
    __shared__ double red[WAVES][2];
    const int lane = threadIdx.x, w = threadIdx.y;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = MAX ? wave_max(v[i]) : wave_sum(v[i]);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) red[w][i] = v[i];
    __syncthreads();
    if (lane == 0 && w == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double a = red[0][i];
            for (int k = 1; k < WAVES; ++k) {
                double b = red[k][i];
                a = MAX ? ((isnan(b) || b > a) ? b : a) : a + b;
            }
            v[i] = a;
        }
    }
    __syncthreads();
}

// Per-workgroup partial of a grid-wide reduction: thread (0,0) stores the
// workgroup's total(s) at partials[i*nblocks + block].  The finishing kernel
// k_finish (one workgroup, launched right after on the same stream) sums the
// partials in a fixed order; the kernel boundary orders the two, so no device
// fence (an L2 write-back on this chip) is needed.
template <int NV, bool MAX>
__device__ __forceinline__ void block_partial(double (&v)[NV], double* partials) {
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.x + blockIdx.y * gridDim.x;
    block_reduce<NV, MAX>(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) partials[i * nb + bid] = v[i];
    }
}

// The finishing kernel's sum of nb partials (k_finish, one value), taken by
// every workgroup of a consumer kernel: the same per-thread strides and the
// same tree, so every workgroup gets k_finish's value bit for bit.  All
// threads of the workgroup must call it; all receive the total.
// thread_sum: this thread's terms j = tid, tid + 256, ... summed in that order,
// the loads of 16 consecutive terms issued together; block_total: the
// workgroup's tree over those sums, returned to every thread.  Split so that a
// kernel needing two totals issues both sets of loads before either tree.
__device__ __forceinline__ double thread_sum(const double* __restrict__ p, int nb) {
    constexpr int UF = 16, ST = WAVE * WAVES;
    const int tid = threadIdx.x + threadIdx.y * WAVE;
    double a = 0.0;
    bool first = true;
    for (int j0 = tid; j0 < nb; j0 += UF * ST) {
        double v[UF];
#pragma unroll
        for (int u = 0; u < UF; ++u) v[u] = p[min(j0 + u * ST, nb - 1)];
#pragma unroll
        for (int u = 0; u < UF; ++u) {
            if (j0 + u * ST >= nb) break;
            if (first) {
                a = v[u];
                first = false;
            } else {
                a = a + v[u];
            }
        }
    }
    return a;
}
__device__ __forceinline__ double block_total(double a) {
    __shared__ double total;
    const int tid = threadIdx.x + threadIdx.y * WAVE;
    double one[1] = {a};
    block_reduce<1, false>(one);
    if (tid == 0) total = one[0];
    __syncthreads();
    const double t = total;
    __syncthreads();  // a later call rewrites total
    return t;
}
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int nb) {
    return block_total(thread_sum(p, nb));
}

__device__ __forceinline__ bool lead_thread() {
    return blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0;
}

// --------------------------------------------------------- field kernels

struct FV {  // named views of Fields
    double *h, *pw, *po, *b, *beta, *abs_ub, *mdot, *shear, *potential, *sensible, *Re, *K, *G, *zb, *zs, *H, *ieb,
        *lambda, *A_visc, *N, *mask;
    double *dhdx, *q_x, *Re_x, *b_x, *ub_x, *taub_x, *dpwdx, *valid_x;
    double *dhdy, *q_y, *Re_y, *b_y, *ub_y, *taub_y, *dpwdy, *valid_y;
    // member-wise from the kernel argument, so the pointers stay in scalar
    // registers (a cast-and-store loop would place the struct in scratch)
    __device__ __forceinline__ explicit FV(const Fields& F)
        : h(F.f[SHK_H]), pw(F.f[SHK_PW]), po(F.f[SHK_PO]), b(F.f[SHK_B]), beta(F.f[SHK_BETA]),
          abs_ub(F.f[SHK_ABS_UB]), mdot(F.f[SHK_MDOT]), shear(F.f[SHK_SHEAR]), potential(F.f[SHK_POTENTIAL]),
          sensible(F.f[SHK_SENSIBLE]), Re(F.f[SHK_RE]), K(F.f[SHK_K]), G(F.f[SHK_G]), zb(F.f[SHK_ZB]),
          zs(F.f[SHK_ZS]), H(F.f[SHK_HICE]), ieb(F.f[SHK_IEB]), lambda(F.f[SHK_LAMBDA]), A_visc(F.f[SHK_A_VISC]),
          N(F.f[SHK_N]), mask(F.f[SHK_MASK]), dhdx(F.f[SHK_DHDX]), q_x(F.f[SHK_Q_X]), Re_x(F.f[SHK_RE_X]),
          b_x(F.f[SHK_B_X]), ub_x(F.f[SHK_UB_X]), taub_x(F.f[SHK_TAUB_X]), dpwdx(F.f[SHK_DPWDX]),
          valid_x(F.f[SHK_VALID_X]), dhdy(F.f[SHK_DHDY]), q_y(F.f[SHK_Q_Y]), Re_y(F.f[SHK_RE_Y]),
          b_y(F.f[SHK_B_Y]), ub_y(F.f[SHK_UB_Y]), taub_y(F.f[SHK_TAUB_Y]), dpwdy(F.f[SHK_DPWDY]),
          valid_y(F.f[SHK_VALID_Y]) {}
};

// Initial conditions, pointwise part (src/initial_conditions.jl:44-96 with
// src/mask.jl:62-150, src/static_fields.jl:8-74, src/gap_height.jl:47-73):
// face masks, lambda, b zeroed off GROUNDED, H, beta, masked ub, abs_ub, po,
// the initial pw (po/2 or the Dirichlet value) and h from pw.
__global__ __launch_bounds__(256) void k_ic_cells(Phys P, Fields F) {
    const int nx = P.nx, ny = P.ny;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    FV s(F);
    const size_t c = cidx(ix, iy, nx);
    const double m = s.mask[c];
    // valid_x on the west face (and the east boundary face), valid_y likewise
    {
        double v = 1.0;
        if (ix > 0) v = (s.mask[c - 1] != OTHER_BASIN && m != OTHER_BASIN) ? 1.0 : 0.0;
        s.valid_x[xidx(ix, iy, nx)] = v;
        if (ix == nx - 1) s.valid_x[xidx(nx, iy, nx)] = 1.0;
        double w = 1.0;
        if (iy > 0) w = (s.mask[c - nx] != OTHER_BASIN && m != OTHER_BASIN) ? 1.0 : 0.0;
        s.valid_y[cidx(ix, iy, nx)] = w;
        if (iy == ny - 1) s.valid_y[cidx(ix, ny, nx)] = 1.0;
    }
    s.lambda[c] = 1.5 * s.A_visc[c];
    double b = s.b[c];
    if (m != GROUNDED) b = 0.0;
    s.b[c] = b;
    const double H = s.zs[c] - (s.zb[c] + b);
    s.H[c] = H;
    if (P.with_open_by_sliding) s.beta[c] = jl_max0((P.br - b) / P.lr);
    // apply_mask_to_sliding!: a face keeps ub only between two GROUNDED cells
    auto ubx_masked = [&](int fx) -> double {
        if (fx == 0 || fx == nx) return 0.0;
        return (s.mask[cidx(fx - 1, iy, nx)] == GROUNDED && s.mask[cidx(fx, iy, nx)] == GROUNDED)
                   ? s.ub_x[xidx(fx, iy, nx)] : 0.0;
    };
    auto uby_masked = [&](int fy) -> double {
        if (fy == 0 || fy == ny) return 0.0;
        return (s.mask[cidx(ix, fy - 1, nx)] == GROUNDED && s.mask[cidx(ix, fy, nx)] == GROUNDED)
                   ? s.ub_y[cidx(ix, fy, nx)] : 0.0;
    };
    const double uw = ubx_masked(ix), ue = ubx_masked(ix + 1), us = uby_masked(iy), un = uby_masked(iy + 1);
    const double ubx = (uw + ue) / 2.0, uby = (us + un) / 2.0;
    s.abs_ub[c] = sqrt(ubx * ubx + uby * uby);
    // Each face is written by exactly one thread (the cell east/north of it;
    // the last column/row also writes the boundary face).  A neighbour may
    // read a face before or after it is overwritten: masking is idempotent
    // (a face keeps its value or becomes 0), so both reads give the same.
    s.ub_x[xidx(ix, iy, nx)] = uw;
    if (ix == nx - 1) s.ub_x[xidx(nx, iy, nx)] = 0.0;
    s.ub_y[cidx(ix, iy, nx)] = us;
    if (iy == ny - 1) s.ub_y[cidx(ix, ny, nx)] = 0.0;
    const double po = P.rho_i * P.g * H;
    s.po[c] = po;
    double pw = po / 2.0;
    if (m == OCEAN) pw = P.p_atm - P.rho_sw * P.g * jl_min(s.zb[c], 0.0);
    else if (m == LAND || m == OTHER_BASIN) pw = P.p_atm;
    s.pw[c] = pw;
    s.h[c] = pw / (P.rho_w * P.g) + s.zb[c];
}

// b_x, b_y from b (src/gap_height.jl:75-115); launched over the union of the
// two face grids, (nx+1) x (ny+1).
__global__ __launch_bounds__(256) void k_stagger_b(Phys P, Fields F) {
    const int nx = P.nx, ny = P.ny;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    FV s(F);
    if (ix <= nx && iy < ny) {
        double v;
        if (ix == 0) v = s.b[cidx(0, iy, nx)];
        else if (ix == nx) v = s.b[cidx(nx - 1, iy, nx)];
        else v = (s.b[cidx(ix, iy, nx)] + s.b[cidx(ix - 1, iy, nx)]) / 2.0;
        s.b_x[xidx(ix, iy, nx)] = v;
    }
    if (ix < nx && iy <= ny) {
        double v;
        if (iy == 0) v = s.b[cidx(ix, 0, nx)];
        else if (iy == ny) v = s.b[cidx(ix, ny - 1, nx)];
        else v = (s.b[cidx(ix, iy, nx)] + s.b[cidx(ix, iy - 1, nx)]) / 2.0;
        s.b_y[cidx(ix, iy, nx)] = v;
    }
}

// Face flux and Reynolds number, exact quadratic (src/water_flux.jl:192-216).
__device__ __forceinline__ void qre(const Phys& P, double bf, double grad, double& q, double& Re) {
    const double G = fabs(grad);
    const double D = (bf * bf * bf * P.g * G) / (12.0 * (P.nu * P.nu));
    const double R = 2.0 * D / (1.0 + sqrt(1.0 + 4.0 * P.omega * D));
    Re = R;
    q = -jl_sign(grad) * P.nu * R;
}

// Face stress of the sliding laws (src/sliding_law.jl): Nf, abs_v and lf are
// the face values of N, |ub| and lambda (outer faces their cell's, inner the
// mean of the two cells), Cf the face coefficient of the field laws.
//   regularised Coulomb (:213-249) and its field form (:360-396):
//     N C (v / (v + |N|^n lambda))^(1/n) ub / v, 0 where v = 0;
//   V0 (:509-543): N C (v / (v + v0))^(1/n) ub / v, 0 where v = 0;
//   linear (:590-602): C^2 N ub, C^2 the square of the staggered C.
__device__ __forceinline__ double taub_face(const Phys& P, int law, double Nf, double abs_v, double lf, double ub,
                                           double Cf) {
    switch (law) {
        case SHK_SLIDING_LINEAR: return (Cf * Cf) * Nf * ub;
        case SHK_SLIDING_REGULARIZED_COULOMB_V0:
            return abs_v > 0.0 ? Nf * Cf * cpow(abs_v / (abs_v + P.v0), P.inv_n_exp) * (ub / abs_v) : 0.0;
        default: {
            const double C = law == SHK_SLIDING_REGULARIZED_COULOMB_FIELD ? Cf : P.C;
            if (abs_v > 0.0)
                return Nf * C * cpow(abs_v / (abs_v + cpow(fabs(Nf), P.n_exp) * lf), P.inv_n_exp) * (ub / abs_v);
            return 0.0;
        }
    }
}

// cells onto faces as the field sliding laws' constructors stagger C
__global__ __launch_bounds__(256) void k_stagger(int nx, int ny, const double* __restrict__ c, double* __restrict__ fx,
                                                 double* __restrict__ fy) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix <= nx && iy < ny) {
        double v;
        if (ix == 0) v = c[cidx(0, iy, nx)];
        else if (ix == nx) v = c[cidx(nx - 1, iy, nx)];
        else v = (c[cidx(ix - 1, iy, nx)] + c[cidx(ix, iy, nx)]) / 2.0;
        fx[xidx(ix, iy, nx)] = v;
    }
    if (ix < nx && iy <= ny) {
        double v;
        if (iy == 0) v = c[cidx(ix, 0, nx)];
        else if (iy == ny) v = c[cidx(ix, ny - 1, nx)];
        else v = (c[cidx(ix, iy - 1, nx)] + c[cidx(ix, iy, nx)]) / 2.0;
        fy[cidx(ix, iy, nx)] = v;
    }
}

// The refresh after a linear solve (the reference runs 9 launches:
// src/elliptic_solver.jl:193-206) in two passes.  k_derived_faces evaluates
// every face once, one thread per (x face, y face) pair of the union grid
// (nx+1) x (ny+1): gradients (interior faces), flux, Reynolds number and the
// sliding stress, written to the face arrays.  k_derived_cells then reads the
// four faces of each cell for Re, the melt-rate terms and K, and writes pw
// and N.  Every value is the same expression of the same inputs as in the
// reference.  mode 1 is the initial-condition variant (pw read as stored
// instead of computed from h, src/initial_conditions.jl:81-124).  With check,
// the cell pass also reduces max|h - h_prev| and max|h| for the Picard test
// (src/elliptic_solver.jl:153-169).
template <int MODE>
__global__ __launch_bounds__(256) void k_derived_faces(Phys P, Fields F) {
    const int nx = P.nx, ny = P.ny;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    FV s(F);
    const double rg = P.rho_w * P.g;
    auto PW = [&](size_t k) -> double { return MODE == 1 ? s.pw[k] : rg * (s.h[k] - s.zb[k]); };
    auto NN = [&](size_t k) -> double { return jl_boolmul(s.mask[k] == GROUNDED, s.po[k] - PW(k)); };
    // x face fx = ix of row iy
    if (ix <= nx && iy < ny) {
        const int fx = ix;
        const size_t k = xidx(fx, iy, nx);
        double dh, dp;
        const bool inner = fx > 0 && fx < nx;
        if (inner) {
            const size_t cw = cidx(fx - 1, iy, nx), ce = cidx(fx, iy, nx);
            const double v = s.valid_x[k];
            dh = ((s.h[ce] - s.h[cw]) / P.dx) * v;
            dp = ((PW(ce) - PW(cw)) / P.dx) * v;
        } else {
            dh = s.dhdx[k];
            dp = s.dpwdx[k];
        }
        double q, Re;
        qre(P, s.b_x[k], dh, q, Re);
        if (inner) {
            s.dhdx[k] = dh;
            s.dpwdx[k] = dp;
        }
        s.q_x[k] = q;
        s.Re_x[k] = Re;
        if (P.sliding_law != SHK_SLIDING_PRESCRIBED) {
            double Nf, av, lf;
            if (fx == 0) { Nf = NN(cidx(0, iy, nx)); av = s.abs_ub[cidx(0, iy, nx)]; lf = s.lambda[cidx(0, iy, nx)]; }
            else if (fx == nx) { const size_t cc = cidx(nx - 1, iy, nx); Nf = NN(cc); av = s.abs_ub[cc]; lf = s.lambda[cc]; }
            else {
                const size_t cw = cidx(fx - 1, iy, nx), ce = cidx(fx, iy, nx);
                Nf = (NN(ce) + NN(cw)) / 2.0;
                lf = (s.lambda[ce] + s.lambda[cw]) / 2.0;
                av = (s.abs_ub[ce] + s.abs_ub[cw]) / 2.0;
            }
            s.taub_x[k] = taub_face(P, P.sliding_law, Nf, av, lf, s.ub_x[k], P.Cx ? P.Cx[k] : 0.0);
        }
    }
    // y face fy = iy of column ix
    if (ix < nx && iy <= ny) {
        const int fy = iy;
        const size_t k = cidx(ix, fy, nx);
        double dh, dp;
        const bool inner = fy > 0 && fy < ny;
        if (inner) {
            const size_t cs = cidx(ix, fy - 1, nx), cn = cidx(ix, fy, nx);
            const double v = s.valid_y[k];
            dh = ((s.h[cn] - s.h[cs]) / P.dy) * v;
            dp = ((PW(cn) - PW(cs)) / P.dy) * v;
        } else {
            dh = s.dhdy[k];
            dp = s.dpwdy[k];
        }
        double q, Re;
        qre(P, s.b_y[k], dh, q, Re);
        if (inner) {
            s.dhdy[k] = dh;
            s.dpwdy[k] = dp;
        }
        s.q_y[k] = q;
        s.Re_y[k] = Re;
        if (P.sliding_law != SHK_SLIDING_PRESCRIBED) {
            double Nf, av, lf;
            if (fy == 0) { const size_t cc = cidx(ix, 0, nx); Nf = NN(cc); av = s.abs_ub[cc]; lf = s.lambda[cc]; }
            else if (fy == ny) { const size_t cc = cidx(ix, ny - 1, nx); Nf = NN(cc); av = s.abs_ub[cc]; lf = s.lambda[cc]; }
            else {
                const size_t cs = cidx(ix, fy - 1, nx), cn = cidx(ix, fy, nx);
                Nf = (NN(cn) + NN(cs)) / 2.0;
                lf = (s.lambda[cn] + s.lambda[cs]) / 2.0;
                av = (s.abs_ub[cn] + s.abs_ub[cs]) / 2.0;
            }
            s.taub_y[k] = taub_face(P, P.sliding_law, Nf, av, lf, s.ub_y[k], P.Cy ? P.Cy[k] : 0.0);
        }
    }
}

template <int MODE, bool CHECK>
__global__ __launch_bounds__(256) void k_derived_cells(Phys P, Fields F, const double* __restrict__ h_prev,
                                                       double* partials) {
    const int nx = P.nx, ny = P.ny;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    FV s(F);
    const double rg = P.rho_w * P.g;
    double dmax = 0.0, hmax = 0.0;
    bool any_nan = false;
    if (ix < nx && iy < ny) {
        const size_t c = cidx(ix, iy, nx);
        const double hc = s.h[c];
        const double pwc = MODE == 1 ? s.pw[c] : rg * (hc - s.zb[c]);
        const double Nc = jl_boolmul(s.mask[c] == GROUNDED, s.po[c] - pwc);
        const size_t kw = xidx(ix, iy, nx), ke = xidx(ix + 1, iy, nx);
        const size_t ks = cidx(ix, iy, nx), kn = cidx(ix, iy + 1, nx);
        const double qx0 = s.q_x[kw], qx1 = s.q_x[ke], qy0 = s.q_y[ks], qy1 = s.q_y[kn];
        const double rex0 = s.Re_x[kw], rex1 = s.Re_x[ke], rey0 = s.Re_y[ks], rey1 = s.Re_y[kn];
        const double dhx0 = s.dhdx[kw], dhx1 = s.dhdx[ke], dhy0 = s.dhdy[ks], dhy1 = s.dhdy[kn];
        const double tbx0 = s.taub_x[kw], tbx1 = s.taub_x[ke], tby0 = s.taub_y[ks], tby1 = s.taub_y[kn];
        const double ubx0 = s.ub_x[kw], ubx1 = s.ub_x[ke], uby0 = s.ub_y[ks], uby1 = s.ub_y[kn];
        if (MODE == 0) s.pw[c] = pwc;
        s.N[c] = Nc;
        // ---- Re (src/water_flux.jl:107-112), mdot (src/melt_rate.jl:106-137), K (:221-226)
        const double Re_c = (rex0 + rex1 + rey0 + rey1) / 4.0;
        s.Re[c] = Re_c;
        const double sh = fabs((ubx1 * tbx1 + ubx0 * tbx0) / 2.0 + (uby1 * tby1 + uby0 * tby0) / 2.0);
        const double pot = fabs((qx1 * dhx1 + qx0 * dhx0) / 2.0 + (qy1 * dhy1 + qy0 * dhy0) / 2.0);
        s.shear[c] = sh;
        s.potential[c] = pot;
        const double Linv = 1.0 / P.L;
        if (P.with_sensible) {
            const double dpx0 = s.dpwdx[kw], dpx1 = s.dpwdx[ke], dpy0 = s.dpwdy[ks], dpy1 = s.dpwdy[kn];
            const double sens = (qx1 * dpx1 + qx0 * dpx0) / 2.0 + (qy1 * dpy1 + qy0 * dpy0) / 2.0;
            s.sensible[c] = sens;
            s.mdot[c] = Linv * (s.G[c] + sh + P.rho_w * P.g * pot + P.ct * P.cw * P.rho_w * sens);
        } else {
            s.mdot[c] = Linv * (s.G[c] + sh + P.rho_w * P.g * pot);
        }
        const double bc = s.b[c];
        s.K[c] = (bc * bc * bc * P.g) / (12.0 * P.nu * (1.0 + P.omega * Re_c));
        if (CHECK) {
            const double d = fabs(hc - h_prev[c]), a = fabs(hc);
            any_nan |= isnan(d) || isnan(a);
            dmax = d > dmax ? d : dmax;
            hmax = a > hmax ? a : hmax;
        }
    }
    if (CHECK) {
        double v[2] = {any_nan ? NAN : dmax, hmax};
        block_partial<2, true>(v, partials);
    }
}

// ------------------------------------------------ one-pass refresh
//
// The refresh of k_derived_faces + k_derived_cells in one marching pass that
// keeps the face values in registers.  Each wave owns DER_OWN = 63 columns
// of a band of R rows.  Per row, lane l evaluates the x face ix = base + l
// (its cell's west face; lane 63's is the east face of lane 62's cell) and
// the y face above its cell; a cell's east face comes from the next lane and
// its south face from the previous row.  Each face is evaluated once, except
// the x faces between strips and the south face of a band's first row, which
// two waves evaluate.  The values are the same expressions of the same inputs
// as in the two-pass kernels, so they agree bit for bit.
// FULL writes every field the two-pass refresh writes.  Otherwise only N, K
// and mdot are written: what the next assembly reads (with the fields a
// Picard loop leaves unchanged) when another Picard iteration follows.

constexpr int DER_OWN = WAVE - 1;

struct DCell {  // what a face reads of an adjacent cell
    double h, pw, N, ub, lam;
};
struct DFace {  // one face: gradients, flux, Reynolds number, stress, |ub| face value
    double dh, dp, q, Re, tb, ub;
};
struct DProd {  // what the cell pass reads of a face: Re, ub*taub, q*dh/dx, q*dpw/dx
    double Re, ut, qh, qp;
};
__device__ __forceinline__ DProd dprod(const DFace& f) {
    DProd p;
    p.Re = f.Re;
    p.ut = f.ub * f.tb;
    p.qh = f.q * f.dh;
    p.qp = f.q * f.dp;
    return p;
}

// LAW: the sliding law as a compile-time constant (the kernel is instantiated
// per law, so each holds only its own stress code), or -1 for P.sliding_law.
template <int LAW>
__device__ __forceinline__ int law_of(const Phys& P) {
    return LAW >= 0 ? LAW : P.sliding_law;
}

template <int MODE, int LAW>
__device__ __forceinline__ DCell dcell(const Phys& P, const FV& s, double rg, size_t c) {
    DCell v;
    v.h = s.h[c];
    v.pw = MODE == 1 ? s.pw[c] : rg * (v.h - s.zb[c]);
    v.N = jl_boolmul(s.mask[c] == GROUNDED, s.po[c] - v.pw);
    const bool slide = law_of<LAW>(P) != SHK_SLIDING_PRESCRIBED;
    v.ub = slide ? s.abs_ub[c] : 0.0;
    v.lam = slide ? s.lambda[c] : 0.0;
    return v;
}

// Face k between cells lo (west or south) and hi (east or north), d the cell
// spacing across it.  edge -1: the domain's first face (only hi exists);
// +1: its last face (only lo exists); 0: an inner face.  Boundary faces keep
// their stored gradients, as in k_derived_faces.
template <int LAW>
__device__ __forceinline__ DFace dface(const Phys& P, int edge, const DCell& lo, const DCell& hi, double d,
                                       const double* __restrict__ dh_a, const double* __restrict__ dp_a,
                                       const double* __restrict__ b_a, const double* __restrict__ ub_a,
                                       const double* __restrict__ v_a, const double* __restrict__ tb_a,
                                       const double* __restrict__ C_a, size_t k) {
    DFace f;
    if (edge == 0) {
        const double v = v_a[k];
        f.dh = ((hi.h - lo.h) / d) * v;
        f.dp = ((hi.pw - lo.pw) / d) * v;
    } else {
        f.dh = dh_a[k];
        f.dp = dp_a[k];
    }
    qre(P, b_a[k], f.dh, f.q, f.Re);
    f.ub = ub_a[k];
    const int law = law_of<LAW>(P);
    if (law != SHK_SLIDING_PRESCRIBED) {
        double Nf, av, lf;
        if (edge < 0) {
            Nf = hi.N; av = hi.ub; lf = hi.lam;
        } else if (edge > 0) {
            Nf = lo.N; av = lo.ub; lf = lo.lam;
        } else {
            Nf = (hi.N + lo.N) / 2.0;
            lf = (hi.lam + lo.lam) / 2.0;
            av = (hi.ub + lo.ub) / 2.0;
        }
        f.tb = taub_face(P, law, Nf, av, lf, f.ub, C_a ? C_a[k] : 0.0);
    } else {
        f.tb = tb_a[k];
    }
    return f;
}

template <int LAW>
__device__ __forceinline__ void store_face(const Phys& P, size_t k, const DFace& f, bool inner, double* dh_a,
                                           double* dp_a, double* q_a, double* Re_a, double* tb_a) {
    if (inner) {
        dh_a[k] = f.dh;
        dp_a[k] = f.dp;
    }
    q_a[k] = f.q;
    Re_a[k] = f.Re;
    if (law_of<LAW>(P) != SHK_SLIDING_PRESCRIBED) tb_a[k] = f.tb;
}

__device__ __forceinline__ DProd prod_from_east(const DProd& f, bool sens) {
    DProd g;
    g.Re = from_east(f.Re);
    g.ut = from_east(f.ut);
    g.qh = from_east(f.qh);
    g.qp = sens ? from_east(f.qp) : 0.0;
    return g;
}
__device__ __forceinline__ DCell cell_from_west(const DCell& c) {
    DCell g;
    g.h = from_west(c.h);
    g.pw = from_west(c.pw);
    g.N = from_west(c.N);
    g.ub = from_west(c.ub);
    g.lam = from_west(c.lam);
    return g;
}

template <int MODE, bool FULL, int LAW>
__global__ __launch_bounds__(256) void k_derived_march(Phys P, Fields F, int R) {
    const int nx = P.nx, ny = P.ny;
    int vbx, vby;
    virtual_block(vbx, vby);
    const int lane = threadIdx.x;
    const int ix = vbx * DER_OWN + lane;
    const int iy0 = (vby * WAVES + threadIdx.y) * R;
    if (iy0 >= ny) return;  // uniform over the wave
    const int iy1 = min(iy0 + R, ny);
    FV s(F);
    const double rg = P.rho_w * P.g;
    const bool own = lane < DER_OWN && ix < nx;                   // this strip's cell
    const bool xstore = ix <= nx && (lane < DER_OWN || ix == nx);  // this strip writes x face ix
    const int ixc = min(ix, nx - 1);                               // clamped column of cell loads
    const int ixf = min(ix, nx);                                   // clamped x face
    const int xedge = ix == 0 ? -1 : (ix >= nx ? 1 : 0);
    const bool sens = P.with_sensible;
    DCell cur = dcell<MODE, LAW>(P, s, rg, cidx(ixc, iy0, nx));
    DProd ps;  // the south face of the current row
    {
        const size_t k = cidx(ixc, iy0, nx);
        DFace fs;
        if (iy0 == 0) {
            fs = dface<LAW>(P, -1, cur, cur, P.dy, s.dhdy, s.dpwdy, s.b_y, s.ub_y, s.valid_y, s.taub_y, P.Cy, k);
            if (FULL && own) store_face<LAW>(P, k, fs, false, s.dhdy, s.dpwdy, s.q_y, s.Re_y, s.taub_y);
        } else {
            const DCell below = dcell<MODE, LAW>(P, s, rg, cidx(ixc, iy0 - 1, nx));
            fs = dface<LAW>(P, 0, below, cur, P.dy, s.dhdy, s.dpwdy, s.b_y, s.ub_y, s.valid_y, s.taub_y, P.Cy, k);
        }
        ps = dprod(fs);
    }
    for (int iy = iy0; iy < iy1; ++iy) {
        const DCell nxt = iy + 1 < ny ? dcell<MODE, LAW>(P, s, rg, cidx(ixc, iy + 1, nx)) : cur;
        DCell west = cell_from_west(cur);
        if (lane == 0 && ix > 0) west = dcell<MODE, LAW>(P, s, rg, cidx(ix - 1, iy, nx));
        const size_t kx = xidx(ixf, iy, nx), ky = cidx(ixc, iy + 1, nx);
        DProd pw_, pe, pn;
        {
            const DFace fw = dface<LAW>(P, xedge, west, cur, P.dx, s.dhdx, s.dpwdx, s.b_x, s.ub_x, s.valid_x, s.taub_x, P.Cx, kx);
            if (FULL && xstore) store_face<LAW>(P, kx, fw, xedge == 0, s.dhdx, s.dpwdx, s.q_x, s.Re_x, s.taub_x);
            pw_ = dprod(fw);
        }
        pe = prod_from_east(pw_, sens);
        const int yedge = iy + 1 == ny ? 1 : 0;
        {
            const DFace fn = dface<LAW>(P, yedge, cur, nxt, P.dy, s.dhdy, s.dpwdy, s.b_y, s.ub_y, s.valid_y, s.taub_y, P.Cy, ky);
            if (FULL && own) store_face<LAW>(P, ky, fn, yedge == 0, s.dhdy, s.dpwdy, s.q_y, s.Re_y, s.taub_y);
            pn = dprod(fn);
        }
        if (own) {
            // the cell pass of k_derived_cells, same expressions
            const size_t c = cidx(ix, iy, nx);
            const double Re_c = (pw_.Re + pe.Re + ps.Re + pn.Re) / 4.0;
            const double sh = fabs((pe.ut + pw_.ut) / 2.0 + (pn.ut + ps.ut) / 2.0);
            const double pot = fabs((pe.qh + pw_.qh) / 2.0 + (pn.qh + ps.qh) / 2.0);
            const double Linv = 1.0 / P.L;
            if (sens) {
                const double sn = (pe.qp + pw_.qp) / 2.0 + (pn.qp + ps.qp) / 2.0;
                if (FULL) s.sensible[c] = sn;
                s.mdot[c] = Linv * (s.G[c] + sh + P.rho_w * P.g * pot + P.ct * P.cw * P.rho_w * sn);
            } else {
                s.mdot[c] = Linv * (s.G[c] + sh + P.rho_w * P.g * pot);
            }
            const double bc = s.b[c];
            s.K[c] = (bc * bc * bc * P.g) / (12.0 * P.nu * (1.0 + P.omega * Re_c));
            s.N[c] = cur.N;
            if (FULL) {
                if (MODE == 0) s.pw[c] = cur.pw;
                s.Re[c] = Re_c;
                s.shear[c] = sh;
                s.potential[c] = pot;
            }
        }
        ps = pn;
        cur = nxt;
    }
}

// The Picard test's maxima (src/elliptic_solver.jl:153-169): max|h - h_prev|
// (NaN if either head holds a NaN) and max|h| per workgroup, finished by
// k_finish (FIN_PICARD), as in k_derived_cells with CHECK.
__global__ __launch_bounds__(256) void k_picard_check(int nx, int ny, const double* __restrict__ h,
                                                      const double* __restrict__ h_prev, double* partials) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    double dmax = 0.0, hmax = 0.0;
    bool any_nan = false;
    if (ix < nx && iy < ny) {
        const size_t c = cidx(ix, iy, nx);
        const double hc = h[c];
        const double d = fabs(hc - h_prev[c]), a = fabs(hc);
        any_nan |= isnan(d) || isnan(a);
        dmax = d > dmax ? d : dmax;
        hmax = a > hmax ? a : hmax;
    }
    double v[2] = {any_nan ? NAN : dmax, hmax};
    block_partial<2, true>(v, partials);
}

// compute_b! (src/gap_height.jl:117-135) then beta (src/gap_height.jl:47-52)
__global__ __launch_bounds__(256) void k_gap(Phys P, Fields F, double dt) {
    const int n = P.nx * P.ny;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FV s(F);
    double b = s.b[i];
    if (s.mask[i] == GROUNDED) {
        const double N = s.N[i];
        const double Np = cpow(fabs(N), P.n1_exp);
        double v;
        if (P.gap_scheme == 0)
            v = (b + dt * (s.mdot[i] / P.rho_i + s.beta[i] * s.abs_ub[i])) / (1.0 + dt * s.A_visc[i] * Np * N);
        else
            v = b + dt * (s.mdot[i] / P.rho_i + s.beta[i] * s.abs_ub[i] - s.A_visc[i] * Np * N * b);
        b = jl_clamp(v, P.b_min, P.b_max);
        s.b[i] = b;
    }
    if (P.with_open_by_sliding) s.beta[i] = jl_max0((P.br - b) / P.lr);
}

__global__ __launch_bounds__(256) void k_fill_ieb(int n, double* ieb, double u, const double* moulin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ieb[i] = moulin ? u + moulin[i] : u;
}

// ------------------------------------------------------------- assembly

// src/linear_solver.jl:13-27
__device__ __forceinline__ double dirichlet_head(const Phys& P, double m, double zb) {
    if (m == OCEAN) {
        const double pw_bc = P.p_atm - P.rho_sw * P.g * jl_min(zb, 0.0);
        return pw_bc / (P.rho_w * P.g) + zb;
    }
    return P.p_atm / (P.rho_w * P.g) + zb;
}

// src/k_face_scheme.jl:32-61
__device__ __forceinline__ double bK(const Phys& P, double K1, double K2, double m2) {
    if (m2 == OTHER_BASIN) return 0.0;
    if (m2 == OCEAN || m2 == LAND) return K1;
    if (P.kfs == 0) return (K1 + K2) / 2.0;
    return (2.0 * K1 * K2) / (K1 + K2 + EPS64);
}

// update_MFLS_kernel! / update_MFLS_parabolic_kernel! (src/linear_solver.jl:438-610)
// fused with the K_face scheme and Dirichlet folding, writing the symmetric
// layout (aP, aE, aN) and rhs.  The Jacobi diagonal is aP itself
// (update_diag_precond!, src/linear_solver.jl:697-700).
// cP (optional) receives the diagonal minus the kept couplings -- the creep
// (or storage) reaction plus the conductances of faces to Dirichlet cells --
// summed without cancellation, for the FP32 flux-form operator.
// With in.r, also the start of a CG solve from x0 = h (k_cg_init): r = rhs -
// A h, x = 0, p = 0.  The row's own aP, aE, aN are at hand, and the couplings
// its west and south neighbours store toward it (aE(ix-1), aN(iy-1)) are the
// same symmetric face expression, so A h is formed with the operator apply's
// products, in its order and under its conditions (march), bit for bit.
// One thread per cell (a plain 2D grid of 64 x 4 workgroups).
template <bool PARABOLIC, class T>
__global__ __launch_bounds__(256) void k_assemble(Phys P, Fields F, double dt, T* __restrict__ aP,
                                                  T* __restrict__ aE, T* __restrict__ aN, T* __restrict__ rhs,
                                                  double* __restrict__ cP, launch::AsmLow lo, launch::AsmInit in) {
    const int nx = P.nx, ny = P.ny;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    FV s(F);
    // the CG start of a row: aPc, aEc, aNc its own, aEw, aNs those its west
    // and south neighbours store toward it
    auto start = [&](size_t c, T aPc, T aEc, T aNc, T aEw, T aNs, T b) {
        if constexpr (sizeof(T) == 8) {  // the FP64 diagonal form only
            if (!in.r) return;
            const double* u = s.h;
            T y = aPc * u[c];
            if (ix < nx - 1) y -= aEc * u[c + 1];
            if (ix > 0) y -= aEw * u[c - 1];
            if (iy < ny - 1) y -= aNc * u[c + nx];
            if (iy > 0) y -= aNs * u[c - nx];
            in.r[c] = b - y;
            in.x[c] = 0;
            in.p[c] = 0;
        }
    };
    {
        const size_t c = cidx(ix, iy, nx);
        // the low-precision copies of one row of the flux form (c, e, n, d)
        auto low = [&](double cc, double ee, double nn, double dd) {
            if (lo.c) {
                const float c32 = (float)cc, e32 = (float)ee, n32 = (float)nn, d32 = (float)dd;
                lo.c[c] = c32;
                lo.e[c] = e32;
                lo.n[c] = n32;
                lo.d[c] = d32;
                if (lo.c16) {
                    lo.c16[c] = bf16_rne(c32);
                    lo.e16[c] = bf16_rne(e32);
                    lo.n16[c] = bf16_rne(n32);
                    lo.d16[c] = bf16_rne(d32);
                }
            }
        };
        const double m = s.mask[c];
        if (m == OCEAN || m == LAND || m == OTHER_BASIN) {
            // an identity row; its neighbours keep no coupling toward it
            const T b = m == OTHER_BASIN ? (T)s.h[c] : (T)dirichlet_head(P, m, s.zb[c]);
            aP[c] = 1.0; aE[c] = 0.0; aN[c] = 0.0;
            rhs[c] = b;
            if (cP) cP[c] = 1.0;
            low(1.0, 0.0, 0.0, 1.0);
            start(c, (T)1.0, (T)0.0, (T)0.0, (T)0.0, (T)0.0, b);
            return;
        }
        const double Kc = s.K[c];
        double mE = 0, mW = 0, mN = 0, mS = 0;
        double fE = 0.0, fW = 0.0, fN = 0.0, fS = 0.0;
        if (ix < nx - 1) { mE = s.mask[c + 1]; fE = bK(P, Kc, s.K[c + 1], mE) / P.dx2; }
        if (ix > 0) { mW = s.mask[c - 1]; fW = bK(P, Kc, s.K[c - 1], mW) / P.dx2; }
        if (iy < ny - 1) { mN = s.mask[c + nx]; fN = bK(P, Kc, s.K[c + nx], mN) / P.dy2; }
        if (iy > 0) { mS = s.mask[c - nx]; fS = bK(P, Kc, s.K[c - nx], mS) / P.dy2; }
        const double Nc = s.N[c], bc = s.b[c], Ac = s.A_visc[c];
        const double Np = cpow(fabs(Nc), P.n1_exp);
        const double diag = PARABOLIC ? (fE + fW + fN + fS) + P.e_v / dt
                                      : (fE + fW + fN + fS) + P.n * P.rho_w * P.g * Ac * Np * bc;
        double drhs = 0.0;
        double cE = 0.0, cN = 0.0;
        if (ix < nx - 1) { if (mE == OCEAN || mE == LAND) drhs += fE * dirichlet_head(P, mE, s.zb[c + 1]); else cE = fE; }
        if (ix > 0) { if (mW == OCEAN || mW == LAND) drhs += fW * dirichlet_head(P, mW, s.zb[c - 1]); }
        if (iy < ny - 1) { if (mN == OCEAN || mN == LAND) drhs += fN * dirichlet_head(P, mN, s.zb[c + nx]); else cN = fN; }
        if (iy > 0) { if (mS == OCEAN || mS == LAND) drhs += fS * dirichlet_head(P, mS, s.zb[c - nx]); }
        // a coupling is kept only between two GROUNDED cells (the neighbour's
        // own row is an identity row otherwise, with its coupling zero)
        if (mE != GROUNDED) cE = 0.0;
        if (mN != GROUNDED) cN = 0.0;
        if (cP) {
            double c0 = PARABOLIC ? P.e_v / dt : P.n * P.rho_w * P.g * Ac * Np * bc;
            if (ix < nx - 1 && (mE == OCEAN || mE == LAND)) c0 += fE;
            if (ix > 0 && (mW == OCEAN || mW == LAND)) c0 += fW;
            if (iy < ny - 1 && (mN == OCEAN || mN == LAND)) c0 += fN;
            if (iy > 0 && (mS == OCEAN || mS == LAND)) c0 += fS;
            cP[c] = c0;
            low(c0, (double)(T)cE, (double)(T)cN, (double)(T)diag);
        }
        double r;
        if (PARABOLIC)
            r = s.mdot[c] * (1.0 / P.rho_w - 1.0 / P.rho_i) - s.beta[c] * s.abs_ub[c] + Ac * Np * Nc * bc + s.ieb[c] +
                (P.e_v / dt) * s.h[c] + drhs;
        else
            r = s.mdot[c] * (1.0 / P.rho_w - 1.0 / P.rho_i) - s.beta[c] * s.abs_ub[c] + Ac * Np * Nc * bc +
                Ac * Np * (P.n * P.rho_w * P.g * s.h[c]) * bc + s.ieb[c] + drhs;
        aP[c] = (T)diag;
        aE[c] = (T)cE;
        aN[c] = (T)cN;
        rhs[c] = (T)r;
        // the west (south) row keeps its coupling toward this one when both
        // are rows of the system and this one is GROUNDED; it is then bK of
        // the same two conductivities, whose form is symmetric in them
        // (sum, and product over sum: bit for bit the same either way round)
        const bool g = m == GROUNDED;
        const bool wRow = !(mW == OCEAN || mW == LAND || mW == OTHER_BASIN);
        const bool sRow = !(mS == OCEAN || mS == LAND || mS == OTHER_BASIN);
        start(c, (T)diag, (T)cE, (T)cN, (g && wRow) ? (T)fW : (T)0.0, (g && sRow) ? (T)fS : (T)0.0, (T)r);
    }
}

// ------------------------------------------------------ marching stencil

// y = A u at every cell of this wave's strip, u given by an operand functor
// (so u can be a stored vector or an expression of several), with the
// epilogue epi(c, Au, u_c) run per cell.  Matches stencil_matvec_kernel!
// (src/linear_solver.jl:616-633): aP*u, then minus the E, W, N, S terms in
// that order, each only where the neighbour exists.  The operand is formed
// once per cell: rows iy-1/iy+1 come from rolling registers, the E/W
// neighbours from wave shuffles (only the two strip-edge lanes evaluate the
// operand of the neighbouring strip).  All 64 lanes stay active so the
// shuffles are well defined; lanes past the east edge only skip their stores.
// FP32 systems are applied in flux form, c*u + sum_f a_f (u - u_f), with aP
// holding c = diagonal minus kept couplings (assembled exactly in FP64): for
// smooth u the FP32 diagonal form would subtract nearly equal numbers.
template <class T, class Opd, class Epi>
__device__ __forceinline__ void march(int nx, int ny, int R, int vbx, int vby, const T* __restrict__ aP,
                                      const T* __restrict__ aE, const T* __restrict__ aN, Opd opd, Epi epi) {
    const int lane = threadIdx.x;
    const int ix = vbx * WAVE + lane;
    const int iy0 = (vby * WAVES + threadIdx.y) * R;
    const int iy1 = min(iy0 + R, ny);
    if (iy0 >= ny) return;  // uniform over the wave
    const bool inx = ix < nx;
    const int ixc = inx ? ix : nx - 1;  // clamped column for loads of idle lanes
    T um = 0, aNm = 0;
    if (iy0 > 0) {
        um = opd(cidx(ixc, iy0 - 1, nx));
        aNm = aN[cidx(ixc, iy0 - 1, nx)];
    }
    T uc = opd(cidx(ixc, iy0, nx));
    T aEw = 0;
#pragma unroll 2
    for (int iy = iy0; iy < iy1; ++iy) {
        const size_t c = cidx(ixc, iy, nx);
        const T up = (iy + 1 < ny) ? opd(c + nx) : T(0);
        const T aPc = aP[c], aEc = aE[c], aNc = aN[c];
        T ue = lane_down1(uc);
        T uw = lane_up1(uc);
        aEw = lane_up1(aEc);
        if (lane == WAVE - 1 && ix < nx - 1) ue = opd(c + 1);
        if (lane == 0 && ix > 0) {
            uw = opd(c - 1);
            aEw = aE[c - 1];
        }
        T y;
        if (sizeof(T) == 4) {
            // FP32: flux form c*u + sum a_f (u - u_nb), aP holding c (no cancellation for smooth u)
            y = aPc * uc;
            if (ix < nx - 1) y += aEc * (uc - ue);
            if (ix > 0) y += aEw * (uc - uw);
            if (iy < ny - 1) y += aNc * (uc - up);
            if (iy > 0) y += aNm * (uc - um);
        } else {
            y = aPc * uc;
            if (ix < nx - 1) y -= aEc * ue;
            if (ix > 0) y -= aEw * uw;
            if (iy < ny - 1) y -= aNc * up;
            if (iy > 0) y -= aNm * um;
        }
        if (inx) epi(c, y, uc);
        um = uc;
        uc = up;
        aNm = aNc;
    }
}

// Two adjacent columns per lane (128-column strips, 16 B loads and stores) for
// the FP64 operator: the same arithmetic in the same order as march(), so
// results are identical.  opd2(c) gives the operand at columns (c, c+1),
// opd1(c) one column (the strip-edge neighbours).  Even nx only (every lane's
// pair is inside the grid or entirely past its east edge).  NT selects
// non-temporal loads and stores for the streamed arrays: grids whose arrays
// exceed the 256 MB Infinity Cache stream them from HBM anyway, and the hint
// keeps them from evicting what is reused.
//
// Tiles (wide_tile): a workgroup's four waves take four adjacent strips over
// the same R rows, so each row of each array is read as 512 contiguous
// columns (4 KB).  From HBM this streams faster than four row blocks of one
// strip: the CG matvec at 4096 x 1024 took 49 against 51 us (57 us for the
// one-column kernel) with its inputs evicted from the caches.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ d2v ld2(const double* p) {
    if constexpr (NT) return __builtin_nontemporal_load((const d2v*)p);
    else return *(const d2v*)p;
}
template <bool NT>
__device__ __forceinline__ void st2(double* p, d2v v) {
    if constexpr (NT) __builtin_nontemporal_store(v, (d2v*)p);
    else *(d2v*)p = v;
}

template <bool NT, class Opd2, class Opd1, class Epi2>
__device__ __forceinline__ void march2(int nx, int ny, int R, int strip, int iy0, const double* __restrict__ aP,
                                       const double* __restrict__ aE, const double* __restrict__ aN, Opd2 opd2,
                                       Opd1 opd1, Epi2 epi2) {
    const int lane = threadIdx.x;
    const int ix = strip * 2 * WAVE + 2 * lane;  // columns ix, ix + 1
    const int iy1 = min(iy0 + R, ny);
    if (iy0 >= ny || strip * 2 * WAVE >= nx) return;  // uniform over the wave
    const bool inx = ix < nx;
    const int ixc = inx ? ix : nx - 2;  // clamped pair for loads of idle lanes
    d2v um = {0.0, 0.0}, aNm = {0.0, 0.0};
    if (iy0 > 0) {
        um = opd2(cidx(ixc, iy0 - 1, nx));
        aNm = ld2<NT>(aN + cidx(ixc, iy0 - 1, nx));
    }
    d2v uc = opd2(cidx(ixc, iy0, nx));
#pragma unroll 2
    for (int iy = iy0; iy < iy1; ++iy) {
        const size_t c = cidx(ixc, iy, nx);
        d2v up = {0.0, 0.0};
        if (iy + 1 < ny) up = opd2(c + nx);
        const d2v aPc = ld2<NT>(aP + c), aEc = ld2<NT>(aE + c), aNc = ld2<NT>(aN + c);
        double ue = lane_down1(uc.x);   // column ix + 2
        double uw = lane_up1(uc.y);     // column ix - 1
        double aEw = lane_up1(aEc.y);   // coupling (ix - 1, ix)
        if (lane == WAVE - 1 && ix + 2 < nx) ue = opd1(c + 2);
        if (lane == 0 && ix > 0) {
            uw = opd1(c - 1);
            aEw = aE[c - 1];
        }
        double y0 = aPc.x * uc.x;
        if (ix < nx - 1) y0 -= aEc.x * uc.y;
        if (ix > 0) y0 -= aEw * uw;
        if (iy < ny - 1) y0 -= aNc.x * up.x;
        if (iy > 0) y0 -= aNm.x * um.x;
        double y1 = aPc.y * uc.y;
        if (ix + 1 < nx - 1) y1 -= aEc.y * ue;
        y1 -= aEc.x * uc.x;
        if (iy < ny - 1) y1 -= aNc.y * up.y;
        if (iy > 0) y1 -= aNm.y * um.y;
        if (inx) epi2(c, d2v{y0, y1}, uc);
        um = uc;
        uc = up;
        aNm = aNc;
    }
}

// Strip and first row of this wave in the wide kernels' tiling (wide_geom).
__device__ __forceinline__ void wide_tile(int R, int& strip, int& iy0) {
    int vbx, vby;
    virtual_block(vbx, vby);
    strip = vbx * WAVES + threadIdx.y;
    iy0 = vby * R;
}

template <bool NT>
__global__ __launch_bounds__(256) void k_matvec2(int nx, int ny, int R, const double* __restrict__ aP,
                                                 const double* __restrict__ aE, const double* __restrict__ aN,
                                                 const double* __restrict__ x, double* __restrict__ y) {
    int strip, iy0;
    wide_tile(R, strip, iy0);
    // the operand is re-read by the strip-edge lanes of the neighbouring strips: plain loads
    march2<NT>(nx, ny, R, strip, iy0, aP, aE, aN, [&](size_t k) -> d2v { return ld2<false>(x + k); },
               [&](size_t k) -> double { return x[k]; }, [&](size_t c, d2v Au, d2v) { st2<NT>(y + c, Au); });
}

// pointwise walk over the same geometry
template <class Epi>
__device__ __forceinline__ void walk(int nx, int ny, int R, int vbx, int vby, Epi epi) {
    const int ix = vbx * WAVE + threadIdx.x;
    const int iy0 = (vby * WAVES + threadIdx.y) * R;
    const int iy1 = min(iy0 + R, ny);
    if (ix >= nx) return;
    for (int iy = iy0; iy < iy1; ++iy) epi(cidx(ix, iy, nx));
}

template <class T>
__global__ __launch_bounds__(256) void k_matvec(int nx, int ny, int R, const T* __restrict__ aP,
                                                const T* __restrict__ aE, const T* __restrict__ aN,
                                                const T* __restrict__ x, T* __restrict__ y) {
    int vbx, vby;
    virtual_block(vbx, vby);
    march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t k) -> T { return x[k]; },
             [&](size_t c, T Au, T) { y[c] = Au; });
}

// ------------------------------------------- Lanczos bound estimate (a9)

// estimate_eigenvalue_bounds (src/preconditioner.jl:89-134): r = rhs, p = r,
// gamma = r'r (x0 = 0).  p_zero is zeroed so step 0 can form p = r + 0*p.
//
// SYM: the same recurrence in the D-inner product <x, y>_D = x'Dy, in which
// D^-1 A is self-adjoint: CG on D^-1 A x = D^-1 rhs from x0 = 0, so r =
// rhs ./ d and gamma = r'Dr; each step's p'Ap is <p, D^-1 A p>_D = p'(A p).
// These are the coefficients of Jacobi-preconditioned CG on A, a symmetric
// Lanczos process, so every Ritz value lies inside the spectrum of D^-1 A
// (the Euclidean dots on the nonsymmetric D^-1 A give no such bound).
template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_lz_init(int nx, int ny, int R, const T* __restrict__ rhs,
                                                 const T* __restrict__ dg, T* __restrict__ r, T* __restrict__ p_zero,
                                                 double* partials, Scalars* sc) {
    int vbx, vby;
    virtual_block(vbx, vby);
    double acc = 0.0;
    walk(nx, ny, R, vbx, vby, [&](size_t c) {
        if constexpr (SYM) {
            const T d = dg[c];
            const T v = rhs[c] / d;
            r[c] = v;
            acc += (double)v * (double)v * (double)d;
        } else {
            const T v = rhs[c];
            r[c] = v;
            acc += (double)v * (double)v;
        }
        p_zero[c] = 0;
    });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// step k: p = r + beta_{k-1} p_old (beta = 0 at k = 0), Ap = (A p) ./ d,
// alpha_k = gamma / p'Ap (SYM: gamma / p'(A p)).
template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_lz_step_a(int nx, int ny, int R, const T* __restrict__ aP,
                                                   const T* __restrict__ aE, const T* __restrict__ aN,
                                                   const T* __restrict__ dg, const T* __restrict__ r,
                                                   const T* __restrict__ p_old, T* __restrict__ p_new,
                                                   T* __restrict__ Ap, int k, double* partials, Scalars* sc) {
    int vbx, vby;
    virtual_block(vbx, vby);
    const T beta = (T)sc->beta;
    double acc = 0.0;
    march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t j) -> T { return r[j] + beta * p_old[j]; },
             [&](size_t c, T Au, T u) {
                 const T a = Au / dg[c];
                 p_new[c] = u;
                 Ap[c] = a;
                 acc += (double)u * (double)(SYM ? Au : a);
             });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// step k (k < nsteps-1): r -= alpha_k Ap; beta_k = r'r / gamma (SYM: r'Dr / gamma).
template <class T, bool SYM>
__global__ __launch_bounds__(256) void k_lz_step_b(int nx, int ny, int R, const T* __restrict__ Ap,
                                                   const T* __restrict__ dg, T* __restrict__ r, int k,
                                                   double* partials, Scalars* sc) {
    int vbx, vby;
    virtual_block(vbx, vby);
    const T alpha = (T)sc->lz_alpha[k];
    double acc = 0.0;
    walk(nx, ny, R, vbx, vby, [&](size_t c) {
        const T v = r[c] - alpha * Ap[c];
        r[c] = v;
        acc += SYM ? (double)v * (double)v * (double)dg[c] : (double)v * (double)v;
    });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// Sturm count: eigenvalues of the tridiagonal (a, e2 = e^2) below x.
__device__ int sturm_count(int m, const double* a, const double* e2, double x) {
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

// One wave: builds the Lanczos tridiagonal from the CG coefficients
// (src/preconditioner.jl:115-123), finds its extreme eigenvalues by Sturm
// multisection (64 points per round; textbook bisection, Golub & Van Loan
// 8.4.1; stands in for LAPACK eigvals(SymTridiagonal) at :125), scales by
// 0.9 / 1.1, and adopts the bounds only if finite, positive and non-empty
// (update_chebyshev_bounds!, :197-233).  Then precomputes this solve's
// Chebyshev alpha/beta sequence (ldiv!, :239-264).
//
// mode 1 keeps the reference's bounds exactly: the recurrence with Euclidean
// dots on the nonsymmetric D^-1 A.  Its Ritz values are not bounded by the
// spectrum: on a 4096 x 1024 glacier with six decades of transmissivity it
// returns lambda_max up to 1e17 (the true value is below 2) and lambda_min
// near 1e-16.  The polynomial then shrinks by that factor, and Krylov's
// zero-curvature test p'Ap <= eps ||p||^2_M, which scales with the
// preconditioner, stops CG early.  Too small a lambda_max is no better: the
// reference's degree-k polynomial turns negative for lambda > lambda_min +
// lambda_max, an indefinite preconditioner.
// mode 0 runs the estimate in the D-inner product (k_lz_* with SYM), whose
// Ritz values lie inside the spectrum, takes lambda_min = 0.9 x the smallest,
// and lambda_max = G, the Gershgorin bound of D^-1 A: every eigenvalue lies
// in [1 - rho, 1 + rho], rho = max_i sum_j |a_ij| / a_ii, so the polynomial is
// positive on the whole spectrum.  The estimate is accepted when finite with
// 0 < lambda_min < G; otherwise the previous lambda_min is kept (the first
// solve keeps the placeholder (1, 1), plain Jacobi sweeps).
__global__ __launch_bounds__(64) void k_cheb_bounds(int nsteps, int degree, int use_estimate, int mode,
                                                    Scalars* sc) {
    __shared__ double a[MAX_LANCZOS], e2[MAX_LANCZOS];
    __shared__ int bad;
    __shared__ double lo_s, hi_s;
    const int t = threadIdx.x;
    if (use_estimate) {
        if (t == 0) {
            bad = 0;
            a[0] = 1.0 / sc->lz_alpha[0];
            for (int k = 1; k < nsteps; ++k) a[k] = 1.0 / sc->lz_alpha[k] + sc->lz_beta[k - 1] / sc->lz_alpha[k - 1];
            double lo = INFINITY, hi = -INFINITY;
            for (int k = 0; k < nsteps; ++k) {
                const double ek = k < nsteps - 1 ? sqrt(sc->lz_beta[k]) / sc->lz_alpha[k] : 0.0;
                if (k < nsteps - 1) {
                    e2[k] = ek * ek;
                    if (!isfinite(ek)) bad = 1;
                }
                if (!isfinite(a[k])) bad = 1;
                double rr = 0.0;
                if (k > 0) rr += fabs(sqrt(sc->lz_beta[k - 1]) / sc->lz_alpha[k - 1]);
                if (k < nsteps - 1) rr += fabs(ek);
                lo = fmin(lo, a[k] - rr);
                hi = fmax(hi, a[k] + rr);
            }
            const double span = fmax(fabs(lo), fabs(hi));
            lo_s = lo - (1e-14 * span + 1e-300);
            hi_s = hi + (1e-14 * span + 1e-300);
        }
        __syncthreads();
        double res[2] = {NAN, NAN};
        if (!bad) {
            for (int which = 0; which < 2; ++which) {
                // smallest: first x with count >= 1; largest: first x with count >= m
                const int target = which == 0 ? 1 : nsteps;
                double lo = lo_s, hi = hi_s;
                for (int round = 0; round < 24; ++round) {
                    const double x = lo + (hi - lo) * (double)(t + 1) / 65.0;
                    const int ok = sturm_count(nsteps, a, e2, x) >= target;
                    // first lane whose point satisfies the count
                    const unsigned long long mask = __ballot(ok);
                    double nlo, nhi;
                    if (mask == 0ull) { nlo = lo + (hi - lo) * 64.0 / 65.0; nhi = hi; }
                    else {
                        const int j = __ffsll((long long)mask) - 1;
                        nlo = j == 0 ? lo : lo + (hi - lo) * (double)j / 65.0;
                        nhi = lo + (hi - lo) * (double)(j + 1) / 65.0;
                    }
                    if (!(nhi - nlo < hi - lo)) break;
                    lo = nlo;
                    hi = nhi;
                }
                res[which] = 0.5 * (lo + hi);
            }
        }
        if (t == 0) {
            const double lmin = res[0] * 0.9, lmax = res[1] * 1.1;
            sc->lz_lmin = lmin;
            sc->lz_lmax = lmax;
            if (mode == 1) {
                if (isfinite(lmin) && isfinite(lmax) && lmin > 0 && lmax > lmin) {
                    sc->lmin = lmin;
                    sc->lmax = lmax;
                    sc->bounds_warning = 0;
                } else {
                    sc->bounds_warning = 1;
                }
            } else {
                const double G = sc->gersh;
                const bool gok = isfinite(G) && G > 0;
                if (gok && isfinite(lmin) && lmin > 0 && G > lmin) {
                    sc->lmin = lmin;
                    sc->lmax = G;
                    sc->bounds_warning = 0;
                } else {
                    sc->bounds_warning = 1;
                    // previous lambda_min with this operator's bound (not the placeholder)
                    if (gok && sc->lmin != sc->lmax && G > sc->lmin) sc->lmax = G;
                }
            }
        }
    }
    if (t == 0) {
        const double d_mid = (sc->lmax + sc->lmin) / 2.0, c_rad = (sc->lmax - sc->lmin) / 2.0;
        double alpha = 1.0 / d_mid;
        sc->cheb_alpha[0] = alpha;
        sc->cheb_beta[0] = 0.0;
        double alpha_prev = alpha;
        for (int k = 1; k < degree; ++k) {
            const double tt = c_rad * alpha_prev / 2.0;
            const double beta = tt * tt;
            alpha = 1.0 / (d_mid - beta / alpha_prev);
            sc->cheb_alpha[k] = alpha;
            sc->cheb_beta[k] = beta;
            alpha_prev = alpha;
        }
        sc->cheb_degree = degree;
    }
}

// ------------------------------------------------------------------ CG

// Finalisers of the preconditioned inner product gamma = r'z.
// mode 0: initial (Krylov cg! set-up), mode 1: per-iteration update.
// A negative or non-finite r'z is a breakdown: the preconditioner (or the
// operator) is not positive definite, and Krylov's rNorm = sqrt(r'z) throws
// (a DomainError in Julia).  The solve stops there and reports it.
__device__ __forceinline__ bool gamma_breaks(double g) { return !(g >= 0.0) || !isfinite(g); }
__device__ __forceinline__ void fin_gamma(Scalars* sc, double g, int mode) {
    if (mode == 0) {
        const double rn = sqrt(g);
        const double ref = sc->rnorm0_ref > 0.0 ? sc->rnorm0_ref : rn;
        sc->gamma = g;
        sc->rNorm0 = ref;
        sc->rNorm = rn;
        sc->pNorm2 = g;
        sc->beta = 0.0;
        sc->iter = 0;
        sc->zero_curv = 0;
        sc->tired = 0;
        sc->eps_stop = sc->atol + sc->rtol * ref;
        sc->pAp = 0.0;
        sc->alpha = 0.0;  // the deferred matvec's first x += alpha p (p = 0) adds nothing
        sc->gpend = 0;
        sc->done_defer = 0;
        sc->breakdown = 0;
        sc->solved = (g == 0.0) || (rn <= sc->eps_stop);
        sc->done = sc->solved;
        if (gamma_breaks(g)) {
            sc->breakdown = 1;
            sc->breakdown_value = g;
            sc->solved = 0;
            sc->done = 1;
        }
    } else {
        if (gamma_breaks(g)) {
            sc->breakdown = 1;
            sc->breakdown_value = g;
            sc->rNorm = sqrt(g);
            sc->solved = 0;
            sc->done = 1;
            return;
        }
        const double rn = sqrt(g);
        sc->rNorm = rn;
        const int solved = (rn <= sc->eps_stop) || (rn + 1.0 <= 1.0);
        sc->solved = solved;
        if (!solved) {
            const double beta = g / sc->gamma;
            sc->pNorm2 = g + beta * beta * sc->pNorm2;
            sc->gamma = g;
            sc->beta = beta;
        }
        sc->iter += 1;
        sc->tired = sc->iter >= sc->itmax;
        sc->done = solved || sc->tired;
    }
}

// Set-up: r = b - A x0 (x0 = h, the warm start; x0 == nullptr means x0 = 0),
// the accumulated correction x = 0, p_zero = 0; with Jacobi also z = r ./ d
// and gamma = r'z.  (Krylov cg! warm start: r = b - A*dx, x = 0, x += dx at exit.)
template <class T>
__global__ __launch_bounds__(256) void k_cg_init(int nx, int ny, int R, const T* __restrict__ aP,
                                                 const T* __restrict__ aE, const T* __restrict__ aN,
                                                 const T* __restrict__ dg, const T* __restrict__ b,
                                                 const T* __restrict__ x0,
                                                 T* __restrict__ r, T* __restrict__ z, T* __restrict__ xacc,
                                                 T* __restrict__ p_zero, int jacobi, double* partials,
                                                 Scalars* sc) {
    int vbx, vby;
    virtual_block(vbx, vby);
    double acc = 0.0;
    if (x0) {
        march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t k) -> T { return x0[k]; },
                 [&](size_t c, T Au, T) {
                     const T rv = b[c] - Au;
                     r[c] = rv;
                     xacc[c] = 0;
                     p_zero[c] = 0;
                     if (jacobi) {
                         const T zv = rv / dg[c];
                         z[c] = zv;
                         acc += (double)rv * (double)zv;
                     }
                 });
    } else {
        walk(nx, ny, R, vbx, vby, [&](size_t c) {
            const T rv = b[c];
            r[c] = rv;
            xacc[c] = 0;
            p_zero[c] = 0;
            if (jacobi) {
                const T zv = rv / dg[c];
                z[c] = zv;
                acc += (double)rv * (double)zv;
            }
        });
    }
    if (jacobi) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
    }
}

// One Chebyshev sweep of ldiv!(y, P, r) (src/preconditioner.jl:239-267),
// pass k = 1 .. degree-1 (the reference's final sweep only updates its
// internal residual, which the returned y never reads, so it is skipped):
//   k == 1: operand p0 = s0 = r ./ d (formed on the fly), then
//           s1 = s0 - a0 (A p0)./d, p1 = s1 + b1 p0, y = a0 p0 + a1 p1
//   k >= 2: s_k = s_{k-1} - a_{k-1} (A p_{k-1})./d, p_k = s_k + b_k p_{k-1},
//           y += a_k p_k
// The last pass (k == degree-1) stores only y and finishes gamma = r'y.
template <class T>
__global__ __launch_bounds__(256) void k_cheb_pass(int nx, int ny, int R, const T* __restrict__ aP,
                                                   const T* __restrict__ aE, const T* __restrict__ aN,
                                                   const T* __restrict__ dg, const T* __restrict__ r,
                                                   const T* __restrict__ s_in,
                                                   const T* __restrict__ pc_in, T* __restrict__ s_out,
                                                   T* __restrict__ pc_out, T* __restrict__ y, int k, int degree,
                                                   int mode, double* partials, Scalars* sc) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    const T a_prev = (T)sc->cheb_alpha[k - 1], a_k = (T)sc->cheb_alpha[k], b_k = (T)sc->cheb_beta[k];
    const bool last = (k == degree - 1);
    double acc = 0.0;
    if (k == 1) {
        march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t j) -> T { return r[j] / dg[j]; },
                 [&](size_t c, T Au, T p0) {
                     const T s1 = p0 - a_prev * (Au / dg[c]);
                     const T p1 = s1 + b_k * p0;
                     const T yv = a_prev * p0 + a_k * p1;
                     if (!last) {
                         s_out[c] = s1;
                         pc_out[c] = p1;
                     } else {
                         acc += (double)r[c] * (double)yv;
                     }
                     y[c] = yv;
                 });
    } else {
        march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t j) -> T { return pc_in[j]; },
                 [&](size_t c, T Au, T pk) {
                     const T sk = s_in[c] - a_prev * (Au / dg[c]);
                     const T pn = sk + b_k * pk;
                     const T yv = y[c] + a_k * pn;
                     if (!last) {
                         s_out[c] = sk;
                         pc_out[c] = pn;
                     } else {
                         acc += (double)r[c] * (double)yv;
                     }
                     y[c] = yv;
                 });
    }
    if (last) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
    }
}

// Chebyshev of degree 1: y = a0 * (r ./ d), gamma = r'y.
template <class T>
__global__ __launch_bounds__(256) void k_cheb_deg1(int nx, int ny, int R, const T* __restrict__ dg,
                                                   const T* __restrict__ r, T* __restrict__ y, int mode,
                                                   double* partials, Scalars* sc) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    const T a0 = (T)sc->cheb_alpha[0];
    double acc = 0.0;
    walk(nx, ny, R, vbx, vby, [&](size_t c) {
        const T rv = r[c];
        const T yv = a0 * (rv / dg[c]);
        y[c] = yv;
        acc += (double)rv * (double)yv;
    });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// p = z + beta*p_old (formed at the stencil points), Ap = A p, p'Ap; then
// Krylov's zero-curvature test and alpha = gamma / p'Ap.
//
// Deferred finish (FP64 CG with the V-cycle, DEFER): the preconditioner of
// the previous iteration left its r'z partials (pa, nba) without a finishing
// kernel (gpend set).  Every workgroup sums them as k_finish would and takes
// the same stop test and beta (Krylov's cg!, as in fin_gamma); workgroup 0
// stores rNorm, solved, iter, tired, beta and pNorm2, and the stop decision in
// done_defer, never in done: the workgroups of this launch test done on
// entry, so a write of done here would let a late workgroup skip its tile.
// The kernel that follows (the update, or the finest down-leg with the update
// fused in) publishes done.  gamma, which the other workgroups read, is also
// stored by that kernel.  Returns 0 (go on), 1 (converged: take the pending
// x += alpha p_old and stop) or 2 (breakdown: stop, x untouched).
template <bool DEFER>
__device__ __forceinline__ int deferred_gamma(Scalars* sc, const double* __restrict__ pa, int nba, double& beta) {
    if (!DEFER || !sc->gpend) return 0;
    const double g = sum_partials(pa, nba);
    const double rn = sqrt(g);
    if (gamma_breaks(g)) {
        if (lead_thread()) {
            sc->rNorm = rn;
            sc->solved = 0;
            sc->breakdown = 1;
            sc->breakdown_value = g;
            sc->done_defer = 1;
        }
        return 2;
    }
    const int solved = (rn <= sc->eps_stop) || (rn + 1.0 <= 1.0);
    const double b = g / sc->gamma;
    if (lead_thread()) {
        sc->rNorm = rn;
        sc->solved = solved;
        if (!solved) {
            sc->pNorm2 = g + b * b * sc->pNorm2;
            sc->beta = b;
        }
        const int it = sc->iter + 1;
        sc->iter = it;
        sc->tired = it >= sc->itmax;
        sc->done_defer = solved || it >= sc->itmax;
    }
    beta = b;
    return solved ? 1 : 0;
}

// The first kernel after a deferred CG matvec publishes the matvec's stop
// decision as done (workgroup 0) and every one of its workgroups returns.
// done_defer was written by the previous launch, so all workgroups of this
// one read the same value.
__device__ __forceinline__ bool publish_deferred_stop(Scalars* sc) {
    if (!sc->done_defer) return false;
    if (lead_thread()) sc->done = 1;
    return true;
}

// Prologue of the kernel that takes the CG update after a deferred matvec
// (k_cg_update<DEFER>, k_cg_update_flat, the finest down-leg with the update
// fused in).  Every workgroup publishes the matvec's stop decision, or sums
// the p'Ap (pb, nbb) and r'z (pa, nba) partials as k_finish would and takes
// Krylov's zero-curvature test and alpha = r'z / p'Ap; workgroup 0 stores
// them.  A non-finite p'Ap is a breakdown.  Returns true when the kernel
// must return (stopped, zero curvature or breakdown).
__device__ __forceinline__ bool deferred_update_prologue(Scalars* sc, const double* __restrict__ pa, int nba,
                                                         const double* __restrict__ pb, int nbb, double& alpha) {
    if (publish_deferred_stop(sc)) return true;
    const double tb = thread_sum(pb, nbb), ta = thread_sum(pa, nba);
    const double pAp = block_total(tb);
    const double g = block_total(ta);
    const double pn2 = sc->pNorm2;
    const bool zc = pAp <= EPS64 * pn2 && fabs(pAp) <= EPS64 * pn2;
    const bool bad = !isfinite(pAp);
    alpha = g / pAp;
    if (lead_thread()) {
        sc->pAp = pAp;
        sc->gamma = g;
        if (bad) {
            sc->breakdown = 2;
            sc->breakdown_value = pAp;
            sc->done = 1;
        } else if (zc) {
            sc->zero_curv = 1;
            sc->done = 1;
        } else {
            sc->alpha = alpha;
        }
    }
    return zc || bad;
}

// DEFER: x (the accumulated correction) takes the previous iteration's
// x += alpha p here, p being p_old, which this kernel reads anyway; the update
// then reads and writes only r.  When the stop test says converged the
// kernel applies that pending update on its tile and stops; a kernel that
// finds done already set does nothing (the update was applied by the kernel
// that set it, or, after a zero curvature, is not applied, as in Krylov).
template <class T, bool DEFER>
__global__ __launch_bounds__(256) void k_cg_matvec(int nx, int ny, int R, const T* __restrict__ aP,
                                                   const T* __restrict__ aE, const T* __restrict__ aN,
                                                   const T* __restrict__ z, const T* __restrict__ p_old,
                                                   T* __restrict__ p_new, T* __restrict__ Ap, double* partials,
                                                   Scalars* sc, const double* __restrict__ pa, int nba,
                                                   T* __restrict__ x) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    double b = sc->beta;
    const T alpha = (T)sc->alpha;  // the previous iteration's step (DEFER)
    if (const int stop = deferred_gamma<DEFER>(sc, pa, nba, b)) {
        if (stop == 1) walk(nx, ny, R, vbx, vby, [&](size_t c) { x[c] = x[c] + alpha * p_old[c]; });
        return;
    }
    const T beta = (T)b;
    double acc = 0.0;
    march<T>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t j) -> T { return z[j] + beta * p_old[j]; },
             [&](size_t c, T Au, T u) {
                 p_new[c] = u;
                 Ap[c] = Au;
                 if constexpr (DEFER) x[c] = x[c] + alpha * p_old[c];
                 acc += (double)u * (double)Au;
             });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// k_cg_matvec on two columns per lane (FP64, even nx).  z is FP64, or the
// V-cycle's FP32 result (ZT = float), widened as it is loaded: the same
// values the FP64 copy would hold.
__device__ __forceinline__ d2v ldz2(const double* p) { return ld2<false>(p); }
__device__ __forceinline__ d2v ldz2(const float* p) {
    const f2v v = *(const f2v*)p;
    return d2v{(double)v.x, (double)v.y};
}
//
// XS (DEFER) selects how the pending x steps are taken.  1: each matvec takes
// the previous iteration's step, x += alpha p_old.  Pairs (0 and 2, in turn):
// the matvec of an odd iteration j takes none and leaves step j-1 owed
// (alpha_pend, xpend_p, xpend_n = 1); the next one, of iteration j+1, takes
// both, x = (x + alpha_pend p_{j-1}) + alpha p_j, reading p_{j-1} from p_new
// (the ping-pong buffer it then overwrites) before its store to each cell.
// The same operations in the same order as two single steps, so the same x,
// bit for bit, with x read and written once per two iterations.  Every end
// of the solve takes the owed step where the single steps would have taken
// it: a stop in the matvec (converged: every pending step; breakdown: the
// owed one only), k_cg_defer_stop likewise, and k_cg_xflush at the end of the
// solve for the rest (tired, zero curvature after an odd iteration).
template <bool NT, bool DEFER, class ZT = double, int XS = 1>
__global__ __launch_bounds__(256) void k_cg_matvec2(int nx, int ny, int R, const double* __restrict__ aP,
                                                    const double* __restrict__ aE, const double* __restrict__ aN,
                                                    const ZT* __restrict__ z, const double* __restrict__ p_old,
                                                    double* p_new, double* __restrict__ Ap,
                                                    double* partials, Scalars* sc, const double* __restrict__ pa,
                                                    int nba, double* __restrict__ x) {
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);
    double beta = sc->beta;
    const double alpha = sc->alpha;  // the previous iteration's step (DEFER), as k_cg_matvec
    const double alpha_pend = XS == 2 ? sc->alpha_pend : 0.0;  // the step owed from the iteration before
    // the pending steps of the pair (XS 2: the owed one from p_new, read
    // before p_new is stored; then the previous iteration's, unless `owed_only`)
    auto x_step = [&](size_t c, d2v pp, bool owed_only) {
        d2v xv = ld2<NT>(x + c);
        if constexpr (XS == 2) {
            xv.x = xv.x + alpha_pend * pp.x;
            xv.y = xv.y + alpha_pend * pp.y;
        }
        if (!owed_only) {
            const d2v po = ld2<false>(p_old + c);
            xv.x = xv.x + alpha * po.x;
            xv.y = xv.y + alpha * po.y;
        }
        st2<NT>(x + c, xv);
    };
    if (const int stop = deferred_gamma<DEFER>(sc, pa, nba, beta)) {
        // converged: march2's cells of this wave, columns ix, ix + 1 of rows iy0 .. iy1 - 1
        const int ix = strip * 2 * WAVE + 2 * (int)threadIdx.x;
        const int iy1 = min(iy0 + R, ny);
        if (ix < nx && (stop == 1 || XS == 2))
            for (int iy = iy0; iy < iy1; ++iy) {
                const size_t c = cidx(ix, iy, nx);
                x_step(c, XS == 2 ? ld2<NT>(p_new + c) : d2v{0.0, 0.0}, stop != 1);
            }
        if constexpr (XS == 2) {
            if (lead_thread()) sc->xpend_n = 0;
        }
        return;
    }
    if constexpr (XS == 0) {
        if (lead_thread()) {
            sc->alpha_pend = alpha;
            sc->xpend_p = p_old;
            sc->xpend_n = 1;
        }
    } else if constexpr (XS == 2) {
        if (lead_thread()) sc->xpend_n = 0;
    }
    double acc = 0.0;
    march2<NT>(
        nx, ny, R, strip, iy0, aP, aE, aN,
        [&](size_t j) -> d2v { return ldz2(z + j) + beta * ld2<false>(p_old + j); },
        [&](size_t j) -> double { return (double)z[j] + beta * p_old[j]; },
        [&](size_t c, d2v Au, d2v u) {
            const d2v pp = XS == 2 ? ld2<NT>(p_new + c) : d2v{0.0, 0.0};
            st2<NT>(p_new + c, u);
            st2<NT>(Ap + c, Au);
            if constexpr (DEFER && XS != 0) x_step(c, pp, false);
            acc += u.x * Au.x;
            acc += u.y * Au.y;
        });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// The end of a batch of deferred CG iterations: the stop decision that the
// next iteration's CG matvec would take (deferred_gamma on the r'z partials
// of the batch's last V-cycle).  When the solve stops there, this kernel makes
// that matvec's writes to the scalars and its pending step x += alpha p_old
// (the same expression, on the flat index), and raises done_defer, on which
// the host stops; a batch then needs no iteration past its last V-cycle.
// When the solve goes on it writes only rNorm (the value the next matvec
// writes), and the next batch's matvec takes the same decision.  It never
// writes done, which every workgroup reads on entry.  (64, 4) workgroups,
// grid-stride over pairs of cells.
// xs: the x-step mode (k_cg_matvec2's XS) of the matvec it stands in for;
// with 2 it also takes the owed step, from p_pend (that matvec's p_new),
// on a stop of either kind.
__global__ __launch_bounds__(256) void k_cg_defer_stop(size_t n, Scalars* sc, const double* __restrict__ pa,
                                                       int nba, const double* __restrict__ p_old,
                                                       double* __restrict__ x, int xs,
                                                       const double* __restrict__ p_pend) {
    if (sc->done || !sc->gpend) return;
    const double g = sum_partials(pa, nba);
    const double rn = sqrt(g);
    // the same tests as deferred_gamma; iter is written below only when every
    // workgroup stops, and then a later reader of it stops as well
    const bool stop = gamma_breaks(g) || rn <= sc->eps_stop || rn + 1.0 <= 1.0 || sc->iter + 1 >= sc->itmax;
    if (!stop) {
        if (lead_thread()) sc->rNorm = rn;
        return;
    }
    const double alpha = sc->alpha;
    const double alpha_pend = sc->alpha_pend;
    double beta;
    // breakdown: no step, as in the matvec (but the owed one)
    const bool owed_only = deferred_gamma<true>(sc, pa, nba, beta) == 2;
    if (xs == 2 && lead_thread()) sc->xpend_n = 0;
    if (owed_only && xs != 2) return;
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    auto step = [&](size_t i) {
        double v = x[i];
        if (xs == 2) v = v + alpha_pend * p_pend[i];
        if (!owed_only) v = v + alpha * p_old[i];
        x[i] = v;
    };
    for (size_t i = 2 * ((size_t)blockIdx.x * WAVE * WAVES + tid); i < n; i += 2 * stride) {
        step(i);
        if (i + 1 < n) step(i + 1);
    }
}

// The end of a solve whose CG matvecs take the x steps in pairs: the step an
// odd iteration's matvec left owed, where no later kernel took it.  Reads
// xpend_n and writes nothing of the scalars (k_reset_scalars clears it).
__global__ __launch_bounds__(256) void k_cg_xflush(size_t n, const Scalars* __restrict__ sc,
                                                   double* __restrict__ x) {
    if (!sc->xpend_n) return;
    const double a = sc->alpha_pend;
    const double* __restrict__ p = sc->xpend_p;
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    for (size_t i = 2 * ((size_t)blockIdx.x * WAVE * WAVES + tid); i < n; i += 2 * stride) {
        x[i] = x[i] + a * p[i];
        if (i + 1 < n) x[i + 1] = x[i + 1] + a * p[i + 1];
    }
}

// x += alpha p; r -= alpha Ap; with Jacobi also z = r ./ d and gamma = r'z.
// DEFER: every workgroup sums p'Ap (pb, nbb) and r'z (pa, nba) as k_finish
// would, applies the zero-curvature test and alpha = r'z / p'Ap (FIN_PAP);
// workgroup 0 stores pAp, gamma, alpha (or zero_curv and done); x += alpha p
// is left to the next CG matvec (k_cg_matvec with DEFER).
template <class T, bool DEFER>
__global__ __launch_bounds__(256) void k_cg_update(int nx, int ny, int R, const T* __restrict__ dg,
                                                   const T* __restrict__ p, const T* __restrict__ Ap,
                                                   T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                   int jacobi, double* partials, Scalars* sc,
                                                   const double* __restrict__ pa, int nba,
                                                   const double* __restrict__ pb, int nbb) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    double al = sc->alpha;
    if constexpr (DEFER) {
        if (deferred_update_prologue(sc, pa, nba, pb, nbb, al)) return;
    }
    const T alpha = (T)al;
    double acc = 0.0;
    walk(nx, ny, R, vbx, vby, [&](size_t c) {
        if constexpr (!DEFER) x[c] = x[c] + alpha * p[c];  // DEFER: in the next CG matvec
        const T rv = r[c] - alpha * Ap[c];
        r[c] = rv;
        if (jacobi) {
            const T zv = rv / dg[c];
            z[c] = zv;
            acc += (double)rv * (double)zv;
        }
    });
    if (jacobi) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
    }
}

// k_cg_update<double, true> (no Jacobi) over the flat cell index: two cells per
// lane (16 B accesses), grid-stride.  Pointwise work streams faster flat than
// in the stencil kernels' strips: 36 against 42 us at 4096 x 1024 with the
// vectors evicted from the caches.  (64, 4) workgroups, as sum_partials needs.
// r -= alpha Ap only: x += alpha p is taken by the next CG matvec.
__global__ __launch_bounds__(256) void k_cg_update_flat(size_t n, const double* __restrict__ Ap,
                                                        double* __restrict__ r, Scalars* sc,
                                                        const double* __restrict__ pa, int nba,
                                                        const double* __restrict__ pb, int nbb) {
    if (sc->done) return;
    double alpha;
    if (deferred_update_prologue(sc, pa, nba, pb, nbb, alpha)) return;
    const size_t tid = threadIdx.x + (size_t)threadIdx.y * WAVE;
    for (size_t i = ((size_t)blockIdx.x * (WAVE * WAVES) + tid) * 2; i < n; i += (size_t)gridDim.x * 2 * WAVE * WAVES) {
        if (i + 1 < n) {
            const d2v av = *(const d2v*)(Ap + i);
            d2v rv = *(const d2v*)(r + i);
            rv.x = rv.x - alpha * av.x;
            rv.y = rv.y - alpha * av.y;
            *(d2v*)(r + i) = rv;
        } else {
            r[i] = r[i] - alpha * Ap[i];
        }
    }
}

// h_new = dx + h_prev (Krylov's warm-start epilogue, then s.h .= ws.x);
// optional under-relaxation h = a h + (1-a) h_prev (src/elliptic_solver.jl:42-46).
__global__ __launch_bounds__(256) void k_cg_finish(int n, const double* __restrict__ x,
                                                   const double* __restrict__ h_prev, double* __restrict__ h_new,
                                                   double relax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double h = x[i] + h_prev[i];
    if (relax > 0.0) h = relax * h + (1.0 - relax) * h_prev[i];
    h_new[i] = h;
}

// Plain streaming copy (reference point for bandwidth and counter scaling).
__global__ __launch_bounds__(256) void k_copy(size_t n, const double* __restrict__ a, double* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Cache flush for cold timings: a grid-stride read of n doubles, one sum per
// workgroup into sink (so the loads cannot be dropped).
__global__ __launch_bounds__(256) void k_read_flush(size_t n, const double* __restrict__ a, double* __restrict__ sink) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) s += a[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < 256; ++k) t += red[k];
        sink[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void k_zero(size_t n, double* p) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------- fused CG update + Chebyshev (H = degree-1)

// One CG iteration's second half in a single pass:
//   x += alpha p;  r_new = r_old - alpha Ap;  z = P^-1 r_new (all degree-1
//   Chebyshev sweeps of ldiv!, src/preconditioner.jl:239-267);  partial of r_new'z.
// Each wave owns a 64-lane strip of columns that overlaps its neighbours by
// H columns on each side; the H sweeps run as a pipeline of stages that
// trail each other by one row (stage k works on row j-k while row j is
// loaded), y-neighbours come from rolling registers and x-neighbours from wave
// shuffles, so the intermediate sweep vectors never touch memory.  A stage
// value is correct on lanes [k, 64-k) and the strip stores lanes [H, 64-H).
// Each quantity is the same expression, in the same order, as in the
// unfused passes (k_cg_update + k_cheb_pass), so the results are identical.
// Reads r_old/Ap/aP/aE/aN with the halo; writes x, r_new, z on owned cells
// only (r is ping-ponged), so no workgroup reads what another writes.

// (Non-temporal hints were measured here as well: at 8192 x 2048 they slowed
// this kernel from 320 to 452 us, the overlapping strips re-reading halo
// columns the hint had let go.)
template <class T, int H>
__global__ __launch_bounds__(256) void k_cg_update_cheb(int nx, int ny, int R, const T* __restrict__ aP,
                                                        const T* __restrict__ aE, const T* __restrict__ aN,
                                                        const T* __restrict__ dg, const T* __restrict__ p,
                                                        const T* __restrict__ Ap,
                                                        const T* __restrict__ r_old, T* __restrict__ r_new,
                                                        T* __restrict__ x, T* __restrict__ z, double* partials,
                                                        Scalars* sc) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    const int lane = threadIdx.x;
    const int ix = vbx * (WAVE - 2 * H) - H + lane;
    const bool inx = ix >= 0 && ix < nx;
    const bool own_x = inx && lane >= H && lane < WAVE - H;
    const bool has_e = ix < nx - 1, has_w = ix > 0;
    const int ixc = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);  // clamped column: every lane loads a valid address
    const int iy0 = (vby * WAVES + threadIdx.y) * R;
    const int iy1 = min(iy0 + R, ny);
    const T alpha = (T)sc->alpha;
    T ca[H + 1], cb[H + 1];
#pragma unroll
    for (int k = 0; k <= H; ++k) {
        ca[k] = (T)sc->cheb_alpha[k];
        cb[k] = (T)sc->cheb_beta[k];
    }
    constexpr bool FLUX = sizeof(T) == 4;  // FP32: flux form, diagonal from dg
    T wP[H + 1], wE[H + 1], wW[H + 1], wN[H + 2], wR[H + 1], wD[H + 1];
    T pw[H][3];
    T sw[H][2], yw[H][2];
#pragma unroll
    for (int k = 0; k <= H; ++k) { wP[k] = 1; wE[k] = 0; wW[k] = 0; wR[k] = 0; wD[k] = 1; }
#pragma unroll
    for (int k = 0; k < H + 2; ++k) wN[k] = 0;
#pragma unroll
    for (int k = 0; k < H; ++k) {
        pw[k][0] = pw[k][1] = pw[k][2] = 0;
        sw[k][0] = sw[k][1] = 0;
        yw[k][0] = yw[k][1] = 0;
    }
    double acc = 0.0;
    const int j0 = max(iy0 - H, 0);
    const int j1 = (iy0 < ny) ? iy1 + H : j0;  // exclusive
    const T* pP = aP + ixc;
    const T* pE = aE + ixc;
    const T* pN = aN + ixc;
    const T* pR = r_old + ixc;
    const T* pA = Ap + ixc;
    const T* pD = dg + ixc;
    auto row_ofs = [&](int j) -> size_t { return (size_t)(j < ny ? j : ny - 1) * (size_t)nx; };
    T nP = 1, nE = 0, nN = 0, nR = 0, nA = 0, nD = 1;  // prefetched inputs of row j
    if (j0 < j1) {
        const size_t o0 = row_ofs(j0);
        nP = pP[o0]; nE = pE[o0]; nN = pN[o0]; nR = pR[o0]; nA = pA[o0];
        if (FLUX) nD = pD[o0];
    }
    // unrolled so that the window shifts become register renames
#pragma unroll 4
    for (int j = j0; j < j1; ++j) {
#pragma unroll
        for (int k = H; k > 0; --k) {
            wP[k] = wP[k - 1]; wE[k] = wE[k - 1]; wW[k] = wW[k - 1]; wR[k] = wR[k - 1]; wD[k] = wD[k - 1];
        }
#pragma unroll
        for (int k = H + 1; k > 0; --k) wN[k] = wN[k - 1];
        // ---- stage 0 at row j (inputs prefetched); issue the loads of row j+1 first
        const bool live = inx && j < ny;
        const T dP = live ? nP : T(1), dE = live ? nE : T(0), dN = live ? nN : T(0);
        const T dD = FLUX ? (live ? nD : T(1)) : dP;
        const T r0 = nR, a0 = nA;
        {
            const size_t o1 = row_ofs(j + 1);
            nP = pP[o1]; nE = pE[o1]; nN = pN[o1]; nR = pR[o1]; nA = pA[o1];
            if (FLUX) nD = pD[o1];
        }
        const T rr = live ? r0 - alpha * a0 : T(0);
        if (own_x && j >= iy0 && j < iy1) {
            const size_t c = cidx(ix, j, nx);
            x[c] = x[c] + alpha * p[c];
            r_new[c] = rr;
        }
        const T pv = rr / dD;
        const T dW = from_west(dE);
        wP[0] = dP; wE[0] = dE; wW[0] = dW; wN[0] = dN; wR[0] = rr; wD[0] = dD;
        T p_in = pv, s_in = pv, y_in = ca[0] * pv;
        // ---- stages 1..H: stage k at row o = j - k
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            pw[k - 1][2] = pw[k - 1][1];
            pw[k - 1][1] = pw[k - 1][0];
            pw[k - 1][0] = p_in;
            sw[k - 1][1] = sw[k - 1][0];
            sw[k - 1][0] = s_in;
            yw[k - 1][1] = yw[k - 1][0];
            yw[k - 1][0] = y_in;
            const int o = j - k;
            const T pc = pw[k - 1][1], pn = pw[k - 1][0], ps = pw[k - 1][2];
            const T pe = from_east(pc);
            const T pwst = from_west(pc);
            // absent neighbours contribute no term (y - 0 == y exactly)
            const T tE = has_e ? wE[k] * pe : T(0);
            const T tW = has_w ? wW[k] * pwst : T(0);
            const T tN = (o < ny - 1) ? wN[k] * pn : T(0);
            const T tS = (o > 0) ? wN[k + 1] * ps : T(0);
            T Au = wP[k] * pc;
            if (FLUX) {
                Au += has_e ? wE[k] * (pc - pe) : T(0);
                Au += has_w ? wW[k] * (pc - pwst) : T(0);
                Au += (o < ny - 1) ? wN[k] * (pc - pn) : T(0);
                Au += (o > 0) ? wN[k + 1] * (pc - ps) : T(0);
            } else {
                Au -= tE;
                Au -= tW;
                Au -= tN;
                Au -= tS;
            }
            const T sk = sw[k - 1][1] - ca[k - 1] * (Au / wD[k]);
            const T pk = sk + cb[k] * pc;
            const T yk = yw[k - 1][1] + ca[k] * pk;
            p_in = pk;
            s_in = sk;
            y_in = yk;
            if (k == H) {
                if (own_x && o >= iy0 && o < iy1) {
                    z[cidx(ix, o, nx)] = yk;
                    acc += (double)wR[k] * (double)yk;
                }
            }
        }
    }
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// --------------------------------------------------- mixed precision pieces

__global__ __launch_bounds__(256) void k_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// FP64 residual of the FP64 system, r32 = (float) r, and the partial of
// r' D^-1 r (D = aP) for the refinement's stopping test.
__global__ __launch_bounds__(256) void k_ir_residual(int nx, int ny, int R, const double* __restrict__ aP,
                                                     const double* __restrict__ aE, const double* __restrict__ aN,
                                                     const double* __restrict__ b, const double* __restrict__ x,
                                                     float* __restrict__ r32, double* partials) {
    int vbx, vby;
    virtual_block(vbx, vby);
    double acc = 0.0;
    march<double>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t k) -> double { return x[k]; },
                  [&](size_t c, double Au, double) {
                      const double r = b[c] - Au;
                      r32[c] = (float)r;
                      acc += r * r / aP[c];
                  });
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// Gershgorin bound of D^-1 A (D = dg): per cell 1 + sum_j |a_ij| / d_i over
// the couplings the operator applies (none past the grid's edges), max over
// the grid (NaN propagates: a zero diagonal with zero couplings gives 0/0).
template <class T>
__global__ __launch_bounds__(256) void k_gersh(int nx, int ny, const T* __restrict__ dg, const T* __restrict__ aE,
                                               const T* __restrict__ aN, double* partials) {
    const int ix = blockIdx.x * WAVE + threadIdx.x, iy = blockIdx.y * WAVES + threadIdx.y;
    double m = 0.0;
    if (ix < nx && iy < ny) {
        const size_t c = cidx(ix, iy, nx);
        double s = 0.0;
        if (ix < nx - 1) s += fabs((double)aE[c]);
        if (ix > 0) s += fabs((double)aE[c - 1]);
        if (iy < ny - 1) s += fabs((double)aN[c]);
        if (iy > 0) s += fabs((double)aN[c - nx]);
        m = 1.0 + s / fabs((double)dg[c]);
    }
    double v[1] = {m};
    block_partial<1, true>(v, partials);
}

// The FP64 residual of a returned head h = x + x0 (x0 may be null), formed
// as k_cg_finish forms it: partials of ||b - A h||^2 and ||b||^2.
// h_out: where given, also the returned head h = x + x0 of every cell (the
// operand at the cell), under-relaxed towards x0 when relax > 0 (as
// k_cg_finish writes it)
// With ppart (and h_out) it also takes the Picard test of the head it writes
// against x0, the previous head (the maxima of k_picard_check, per workgroup).
__global__ __launch_bounds__(256) void k_true_resid(int nx, int ny, int R, const double* __restrict__ aP,
                                                    const double* __restrict__ aE, const double* __restrict__ aN,
                                                    const double* __restrict__ b, const double* __restrict__ x,
                                                    const double* __restrict__ x0, double* partials,
                                                    double* __restrict__ h_out, double relax, double* ppart) {
    int vbx, vby;
    virtual_block(vbx, vby);
    double ar = 0.0, ab = 0.0, dmax = 0.0, hmax = 0.0;
    bool any_nan = false;
    march<double>(nx, ny, R, vbx, vby, aP, aE, aN, [&](size_t k) -> double { return x0 ? x[k] + x0[k] : x[k]; },
                  [&](size_t c, double Au, double u) {
                      const double bc = b[c];
                      const double r = bc - Au;
                      ar += r * r;
                      ab += bc * bc;
                      if (h_out) {
                          const double hn = relax > 0.0 ? relax * u + (1.0 - relax) * x0[c] : u;
                          h_out[c] = hn;
                          if (ppart) {
                              const double d = fabs(hn - x0[c]), a = fabs(hn);
                              any_nan |= isnan(d) || isnan(a);
                              dmax = d > dmax ? d : dmax;
                              hmax = a > hmax ? a : hmax;
                          }
                      }
                  });
    double v[2] = {ar, ab};
    block_partial<2, false>(v, partials);
    if (ppart) {
        double w[2] = {any_nan ? NAN : dmax, hmax};
        block_partial<2, true>(w, ppart);
    }
}

// k_true_resid on two columns per lane (march2, the wide tiling; even nx):
// the same operand, A h and per-cell expressions, so the same head and
// Picard maxima; the norms' partials sum the cells in another order.
__global__ __launch_bounds__(256) void k_true_resid2(int nx, int ny, int R, const double* __restrict__ aP,
                                                     const double* __restrict__ aE, const double* __restrict__ aN,
                                                     const double* __restrict__ b, const double* __restrict__ x,
                                                     const double* __restrict__ x0, double* partials,
                                                     double* __restrict__ h_out, double relax, double* ppart) {
    int strip, iy0;
    wide_tile(R, strip, iy0);
    double ar = 0.0, ab = 0.0, dmax = 0.0, hmax = 0.0;
    bool any_nan = false;
    auto cell = [&](size_t c, double Au, double u) {
        const double bc = b[c];
        const double r = bc - Au;
        ar += r * r;
        ab += bc * bc;
        if (h_out) {
            const double hn = relax > 0.0 ? relax * u + (1.0 - relax) * x0[c] : u;
            h_out[c] = hn;
            if (ppart) {
                const double d = fabs(hn - x0[c]), a = fabs(hn);
                any_nan |= isnan(d) || isnan(a);
                dmax = d > dmax ? d : dmax;
                hmax = a > hmax ? a : hmax;
            }
        }
    };
    march2<false>(
        nx, ny, R, strip, iy0, aP, aE, aN,
        [&](size_t k) -> d2v {
            d2v v = ld2<false>(x + k);
            if (x0) {
                const d2v w = ld2<false>(x0 + k);
                v.x = v.x + w.x;
                v.y = v.y + w.y;
            }
            return v;
        },
        [&](size_t k) -> double { return x0 ? x[k] + x0[k] : x[k]; },
        [&](size_t c, d2v Au, d2v u) {
            cell(c, Au.x, u.x);
            cell(c + 1, Au.y, u.y);
        });
    double v[2] = {ar, ab};
    block_partial<2, false>(v, partials);
    if (ppart) {
        double w[2] = {any_nan ? NAN : dmax, hmax};
        block_partial<2, true>(w, ppart);
    }
}

__global__ __launch_bounds__(256) void k_ir_correct(size_t n, double* __restrict__ x, const float* __restrict__ d) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + (double)d[i];
}

__global__ __launch_bounds__(256) void k_relax(size_t n, double* __restrict__ h, const double* __restrict__ hp,
                                               double a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = a * h[i] + (1.0 - a) * hp[i];
}

// ------------------------------------------------- geometric multigrid (FP32)

// Cell-centred multigrid on the flux-form operator A u = c u + sum_f a_f (u - u_f).
// Coarse cell (I,J) aggregates fine cells (2I..2I+1, 2J..2J+1); with
// piecewise-constant prolongation P and restriction P^T the Galerkin operator
// P^T A P is again a 5-point flux-form operator: c sums over the block, a
// coarse face sums the fine faces it replaces, faces inside the block cancel.

__global__ __launch_bounds__(256) void k_mg_coarsen(int nxf, int nyf, int nxc, int nyc,
                                                    const float* __restrict__ cf, const float* __restrict__ ef,
                                                    const float* __restrict__ nf, float* __restrict__ cc,
                                                    float* __restrict__ ec, float* __restrict__ nc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= nxc || J >= nyc) return;
    const int x0 = 2 * I, y0 = 2 * J;
    const int x1 = min(x0 + 1, nxf - 1), y1 = min(y0 + 1, nyf - 1);
    float c = 0.f, e = 0.f, n = 0.f;
    for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) c += cf[(size_t)x + (size_t)y * nxf];
    for (int y = y0; y <= y1; ++y) e += ef[(size_t)x1 + (size_t)y * nxf];  // fine faces on the block's east side
    for (int x = x0; x <= x1; ++x) n += nf[(size_t)x + (size_t)y1 * nxf];  // ... and north side
    cc[(size_t)I + (size_t)J * nxc] = c;
    ec[(size_t)I + (size_t)J * nxc] = (I < nxc - 1) ? e : 0.f;
    nc[(size_t)I + (size_t)J * nxc] = (J < nyc - 1) ? n : 0.f;
}

// Cell-centred linear transfers.  P gives fine cell x (in 1D) 3/4 of its own
// coarse cell I0 = x / 2 and 1/4 of the coarse neighbour on its side
// (I0 - 1 for even x, I0 + 1 for odd x); a neighbour outside the coarse grid
// folds onto I0 (weight 1).  In 2D the weights multiply (9/16, 3/16, 3/16,
// 1/16).  Restriction is P^T: coarse cell I gathers fine cells 2I-1 .. 2I+2
// with weights 1/4, 3/4, 3/4, 1/4, where fine cell 0 (for I = 0) and an odd
// last fine cell f - 1 (for I = (f+1)/2 - 1) carry weight 1 and cells outside
// the grid none.
__device__ __forceinline__ float lin_rw(int x, int q, int f) {  // fine x = 2I - 1 + q
    if (x < 0 || x >= f) return 0.f;
    if (q == 1 || q == 2) return (x == 0 || (x == f - 1 && (x & 1))) ? 1.f : 0.75f;
    return 0.25f;
}
// Prolongation weights of fine x: its own coarse cell (w0) and the neighbour
// on its side (w1 = 1 - w0, at index I1, folded into the grid).
struct LinP {
    int I0, I1;
    float w0, w1;
};
__device__ __forceinline__ LinP lin_p(int x, int f) {
    const int c = (f + 1) >> 1, I0 = x >> 1, I1 = (x & 1) ? I0 + 1 : I0 - 1;
    const bool out = I1 < 0 || I1 > c - 1;
    return LinP{I0, out ? I0 : I1, out ? 1.f : 0.75f, out ? 0.f : 0.25f};
}

// The coarse operator of the linear transfers: the Galerkin product P^T A P of
// the flux-form operator, lumped back to the 5-point flux form.  c = P^T c;
// a fine face of coupling a between cells i and j adds a (P_i - P_j)(P_i - P_j)^T,
// whose links are laid onto the lattice paths between their end cells (half
// on each of a diagonal link's two paths).  For a fine x-face the links along
// y cancel against the diagonal halves and what remains lies on the coarse
// x-faces of the two coarse rows it touches, weighted by the restriction
// weights in y:
//   ec(I, J) = sum_y wy(y, J) [a(2I+1, y) / 4 + (a(2I, y) + a(2I+2, y)) / 16]
// (the face inside coarse cell I+1 or I reaches across the face I | I+1 with
// a quarter of the quarter-weighted difference of its neighbours' values);
// y-faces likewise.  Sums in FP64, one rounding to FP32.
__global__ __launch_bounds__(256) void k_mg_coarsen_lin(int nxf, int nyf, int nxc, int nyc,
                                                        const float* __restrict__ cf, const float* __restrict__ ef,
                                                        const float* __restrict__ nf, float* __restrict__ cc,
                                                        float* __restrict__ ec, float* __restrict__ nc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= nxc || J >= nyc) return;
    double wx[4], wy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        wx[q] = lin_rw(2 * I - 1 + q, q, nxf);
        wy[q] = lin_rw(2 * J - 1 + q, q, nyf);
    }
    auto E = [&](int x, int y) -> double { return x < nxf - 1 ? (double)ef[(size_t)x + (size_t)y * nxf] : 0.0; };
    auto N = [&](int x, int y) -> double { return y < nyf - 1 ? (double)nf[(size_t)x + (size_t)y * nxf] : 0.0; };
    double c = 0.0, e = 0.0, n = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = 2 * J - 1 + q;
        if (wy[q] == 0.0) continue;  // outside the grid
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int x = 2 * I - 1 + p;
            if (wx[p] != 0.0) s += wx[p] * (double)cf[(size_t)x + (size_t)y * nxf];
        }
        c += wy[q] * s;
        if (I < nxc - 1) e += wy[q] * (0.25 * E(2 * I + 1, y) + 0.0625 * (E(2 * I, y) + E(2 * I + 2, y)));
    }
    if (J < nyc - 1)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int x = 2 * I - 1 + p;
            if (wx[p] != 0.0) n += wx[p] * (0.25 * N(x, 2 * J + 1) + 0.0625 * (N(x, 2 * J) + N(x, 2 * J + 2)));
        }
    const size_t k = (size_t)I + (size_t)J * nxc;
    cc[k] = (float)c;
    ec[k] = (float)e;
    nc[k] = (float)n;
}

// A coefficient as FP32 from FP32 or BF16 storage.
__device__ __forceinline__ float ldc1(const float* p, size_t k) { return p[k]; }
__device__ __forceinline__ float ldc1(const uint16_t* p, size_t k) { return __uint_as_float((uint32_t)p[k] << 16); }

// Restriction by P^T of the residual b - A x of a level whose x is stored:
// one thread per coarse cell, the 4 x 4 fine residuals it gathers recomputed
// from x (each fine residual is taken by up to four coarse cells).  Rows are
// combined with the x weights first, then the rows with the y weights.
template <class TB, class CT>
__global__ __launch_bounds__(256) void k_mg_restrict_lin(int nx, int ny, int nxc, int nyc, const CT* __restrict__ c,
                                                         const CT* __restrict__ e, const CT* __restrict__ n,
                                                         const TB* __restrict__ b, const float* __restrict__ x,
                                                         float* __restrict__ bc, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= nxc || J >= nyc) return;
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int iy = 2 * J - 1 + q;
        const float wy = lin_rw(iy, q, ny);
        if (wy == 0.f) continue;
        float h = 0.f;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int ix = 2 * I - 1 + p;
            const float wx = lin_rw(ix, p, nx);
            if (wx == 0.f) continue;
            const size_t k = (size_t)ix + (size_t)iy * nx;
            const float u = x[k];
            float a = ldc1(c, k) * u;
            if (ix < nx - 1) a += ldc1(e, k) * (u - x[k + 1]);
            if (ix > 0) a += ldc1(e, k - 1) * (u - x[k - 1]);
            if (iy < ny - 1) a += ldc1(n, k) * (u - x[k + nx]);
            if (iy > 0) a += ldc1(n, k - nx) * (u - x[k - nx]);
            h += wx * ((float)b[k] - a);
        }
        acc += wy * h;
    }
    bc[(size_t)I + (size_t)J * nxc] = acc;
}

// u = x + scale P xc in place, one thread per fine cell.
__global__ __launch_bounds__(256) void k_mg_prolong_lin(int nx, int ny, int nxc, float scale, float* __restrict__ x,
                                                        const float* __restrict__ xc, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    const LinP px = lin_p(ix, nx), py = lin_p(iy, ny);
    auto X = [&](int I, int J) -> float { return xc[(size_t)I + (size_t)J * nxc]; };
    const float v0 = px.w0 * X(px.I0, py.I0) + px.w1 * X(px.I1, py.I0);
    const float v1 = px.w0 * X(px.I0, py.I1) + px.w1 * X(px.I1, py.I1);
    const size_t k = (size_t)ix + (size_t)iy * nx;
    x[k] = x[k] + scale * (py.w0 * v0 + py.w1 * v1);
}

// Jacobi diagonal of a flux-form level: c + aE + aW + aN + aS.
__global__ __launch_bounds__(256) void k_mg_diag(int nx, int ny, const float* __restrict__ c,
                                                 const float* __restrict__ e, const float* __restrict__ n,
                                                 float* __restrict__ D) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    const size_t k = (size_t)ix + (size_t)iy * nx;
    float d = c[k] + e[k] + n[k];
    if (ix > 0) d += e[k - 1];
    if (iy > 0) d += n[k - nx];
    D[k] = d;
}

// Flux-form residual helper: (A u)_k with u given by a functor of the index.
template <class U>
__device__ __forceinline__ float flux_apply(int nx, int ny, int ix, int iy, size_t k, const float* __restrict__ c,
                                            const float* __restrict__ e, const float* __restrict__ n, U u) {
    const float uc = u(k);
    float y = c[k] * uc;
    if (ix < nx - 1) y += e[k] * (uc - u(k + 1));
    if (ix > 0) y += e[k - 1] * (uc - u(k - 1));
    if (iy < ny - 1) y += n[k] * (uc - u(k + nx));
    if (iy > 0) y += n[k - nx] * (uc - u(k - nx));
    return y;
}

// FP32 divide through the hardware reciprocal (about 1 ulp) for the V-cycle's
// Jacobi sweeps: the V-cycle only preconditions CG, so its rounding changes
// how fast CG converges, never the solution CG converges to.
__device__ __forceinline__ float mg_div(float a, float d) { return a * __builtin_amdgcn_rcpf(d); }

// ---- red-black Gauss-Seidel smoothing (the default V-cycle smoother) ----
// A cell (ix, iy) is red when ix + iy is even.  Down-leg from x = 0: red
// cells take x = omega b / D, then black cells x = omega (b + sum_f a_f x_f) / D
// from their (red) neighbours, b - A x at a black cell with x_b = 0 being
// b + sum_f a_f x_f in the flux form.  Up-leg, the adjoint order: black cells
// u += omega (b - A u) / D, then red cells the same with the updated black
// neighbours.  Per cell the terms are taken in a fixed order (W, E, N, S for
// the black down-leg value, E, W, N, S for the flux form), and every kernel
// forms a value through these helpers, so a lane that recomputes a
// neighbouring strip's cell gets the owner's bits.  An absent face passes
// zero coupling and zero value.
__device__ __forceinline__ float rb_black0(float b, float d, float omega, float aW, float xW, float aE, float xE,
                                           float aN, float xN, float aS, float xS) {
    float s = b;
    s += aW * xW;
    s += aE * xE;
    s += aN * xN;
    s += aS * xS;
    return mg_div(omega * s, d);
}
// u + omega (b - A u) / D with A u in the flux form (terms E, W, N, S)
__device__ __forceinline__ float rb_relax(float b, float c, float d, float omega, float u, bool hE, float aE,
                                          float uE, bool hW, float aW, float uW, bool hN, float aN, float uN, bool hS,
                                          float aS, float uS) {
    float a = c * u;
    if (hE) a += aE * (u - uE);
    if (hW) a += aW * (u - uW);
    if (hN) a += aN * (u - uN);
    if (hS) a += aS * (u - uS);
    return u + mg_div(omega * (b - a), d);
}
// The down-leg's x at any cell, from global memory (strip-edge lanes, flat kernels).
template <class TB>
__device__ __forceinline__ float rb_down_x(int nx, int ny, int jx, int jy, float omega, const float* __restrict__ e,
                                           const float* __restrict__ n, const float* __restrict__ D,
                                           const TB* __restrict__ b) {
    const size_t k = (size_t)jx + (size_t)jy * nx;
    if (((jx + jy) & 1) == 0) return mg_div(omega * (float)b[k], D[k]);
    const bool hW = jx > 0, hE = jx < nx - 1, hN = jy < ny - 1, hS = jy > 0;
    auto red = [&](size_t q) -> float { return mg_div(omega * (float)b[q], D[q]); };
    return rb_black0((float)b[k], D[k], omega, hW ? e[k - 1] : 0.f, hW ? red(k - 1) : 0.f, hE ? e[k] : 0.f,
                     hE ? red(k + 1) : 0.f, hN ? n[k] : 0.f, hN ? red(k + nx) : 0.f, hS ? n[k - nx] : 0.f,
                     hS ? red(k - nx) : 0.f);
}
// The up-leg's black-updated value at a black cell, u given by a functor of (jx, jy).
template <class TB, class U>
__device__ __forceinline__ float rb_up_black(int nx, int ny, int jx, int jy, float omega, const float* __restrict__ c,
                                             const float* __restrict__ e, const float* __restrict__ n,
                                             const float* __restrict__ D, const TB* __restrict__ b, U u) {
    const size_t k = (size_t)jx + (size_t)jy * nx;
    const bool hE = jx < nx - 1, hW = jx > 0, hN = jy < ny - 1, hS = jy > 0;
    return rb_relax((float)b[k], c[k], D[k], omega, u(jx, jy), hE, hE ? e[k] : 0.f, hE ? u(jx + 1, jy) : 0.f, hW,
                    hW ? e[k - 1] : 0.f, hW ? u(jx - 1, jy) : 0.f, hN, hN ? n[k] : 0.f, hN ? u(jx, jy + 1) : 0.f,
                    hS, hS ? n[k - nx] : 0.f, hS ? u(jx, jy - 1) : 0.f);
}

// The two legs for the large levels as marching kernels (the layout of the
// CG kernels: one wave per 64-column strip walking R rows, R even so a wave
// owns whole row pairs; rows iy-1, iy, iy+1 in rolling registers, E/W
// neighbours by lane shuffles, the strip-edge lanes evaluating the
// neighbouring strip's value).  Every array is read once per launch plus a
// halo row per R rows; the arithmetic is that of the single-phase kernels
// below, in the same order.
//
// Down-leg: x = omega b / D, r = b - A x; the restriction sums the 2 x 2
// residuals as (r00 + r01) + (r10 + r11): horizontal pairs by a shuffle from
// the odd lane, the even row's pair kept in a register until the odd row.
template <class TB>
__global__ __launch_bounds__(256) void k_mg_down_march(int nx, int ny, int nxc, int R, float omega,
                                                       const float* __restrict__ c, const float* __restrict__ e,
                                                       const float* __restrict__ n, const float* __restrict__ D,
                                                       const TB* __restrict__ b, float* __restrict__ x,
                                                       float* __restrict__ bc, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    const int lane = threadIdx.x;
    const int ix = vbx * WAVE + lane;
    const int iy0 = (vby * WAVES + threadIdx.y) * R;  // even
    const int iy1 = min(iy0 + R, ny);
    if (iy0 >= ny) return;  // uniform over the wave; no barriers follow
    const bool inx = ix < nx;
    const int ixc = inx ? ix : nx - 1;  // clamped column for the loads of idle lanes
    const bool even = (lane & 1) == 0;
    auto xs = [&](size_t k) -> float { return mg_div(omega * (float)b[k], D[k]); };
    float um = 0.f, nm = 0.f;
    if (iy0 > 0) {
        um = xs(cidx(ixc, iy0 - 1, nx));
        nm = n[cidx(ixc, iy0 - 1, nx)];
    }
    float uc = xs(cidx(ixc, iy0, nx));
    float hrow = 0.f;  // the even row's horizontal pair sum
    for (int iy = iy0; iy < iy1; ++iy) {
        const size_t k = cidx(ixc, iy, nx);
        const float up = (iy + 1 < ny) ? xs(k + nx) : 0.f;
        const float cc = c[k], ec = e[k], nc = n[k];
        const float bk = (float)b[k];
        float ue = lane_down1(uc);
        float uw = lane_up1(uc);
        float ew = lane_up1(ec);
        if (lane == WAVE - 1 && ix < nx - 1) ue = xs(k + 1);
        if (lane == 0 && ix > 0) {
            uw = xs(k - 1);
            ew = e[k - 1];
        }
        float a = cc * uc;
        if (ix < nx - 1) a += ec * (uc - ue);
        if (ix > 0) a += ew * (uc - uw);
        if (iy < ny - 1) a += nc * (uc - up);
        if (iy > 0) a += nm * (uc - um);
        const float res = inx ? bk - a : 0.f;
        if (inx) x[k] = uc;
        const float hs = res + lane_down1(res);  // meaningful on even lanes
        if ((iy & 1) == 0) {
            hrow = hs;
        } else if (even && inx) {
            bc[(size_t)(ix >> 1) + (size_t)(iy >> 1) * nxc] = hrow + hs;
        }
        um = uc;
        uc = up;
        nm = nc;
    }
    // odd ny: the last (even) row has no partner row; its coarse cells take
    // the even row's sums plus the zero residuals of the missing row
    if ((iy1 & 1) == 1 && even && inx) bc[(size_t)(ix >> 1) + (size_t)((iy1 - 1) >> 1) * nxc] = hrow + 0.f;
}

// Up-leg: u = x + scale P xc, z = u + omega (b - A u) / D (written to out);
// with_dot: the partial of b'z (2: left pending for the next CG matvec).
template <class TB>
__global__ __launch_bounds__(256) void k_mg_up_march(int nx, int ny, int nxc, int R, float omega, float scale,
                                                     const float* __restrict__ c, const float* __restrict__ e,
                                                     const float* __restrict__ n, const float* __restrict__ D,
                                                     const TB* __restrict__ b, const float* __restrict__ x,
                                                     const float* __restrict__ xc, TB* __restrict__ out,
                                                     int with_dot, double* partials, Scalars* __restrict__ sc) {
    if (sc->done) return;
    int vbx, vby;
    virtual_block(vbx, vby);
    const int lane = threadIdx.x;
    const int ix = vbx * WAVE + lane;
    const int iy0 = (vby * WAVES + threadIdx.y) * R;
    const int iy1 = min(iy0 + R, ny);
    double acc = 0.0;
    if (iy0 < ny) {  // uniform over the wave (the partial below needs every wave)
        const bool inx = ix < nx;
        const int ixc = inx ? ix : nx - 1;
        auto us = [&](int jx, int jy) -> float {
            return x[(size_t)jx + (size_t)jy * nx] + scale * xc[(size_t)(jx >> 1) + (size_t)(jy >> 1) * nxc];
        };
        float um = 0.f, nm = 0.f;
        if (iy0 > 0) {
            um = us(ixc, iy0 - 1);
            nm = n[cidx(ixc, iy0 - 1, nx)];
        }
        float uc = us(ixc, iy0);
        for (int iy = iy0; iy < iy1; ++iy) {
            const size_t k = cidx(ixc, iy, nx);
            const float up = (iy + 1 < ny) ? us(ixc, iy + 1) : 0.f;
            const float cc = c[k], ec = e[k], nc = n[k], dk = D[k];
            const TB bk = b[k];
            float ue = lane_down1(uc);
            float uw = lane_up1(uc);
            float ew = lane_up1(ec);
            if (lane == WAVE - 1 && ix < nx - 1) ue = us(ix + 1, iy);
            if (lane == 0 && ix > 0) {
                uw = us(ix - 1, iy);
                ew = e[k - 1];
            }
            float a = cc * uc;
            if (ix < nx - 1) a += ec * (uc - ue);
            if (ix > 0) a += ew * (uc - uw);
            if (iy < ny - 1) a += nc * (uc - up);
            if (iy > 0) a += nm * (uc - um);
            const float z = uc + mg_div(omega * ((float)bk - a), dk);
            if (inx) {
                out[k] = (TB)z;
                acc += (double)bk * (double)z;
            }
            um = uc;
            uc = up;
            nm = nc;
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// The two legs on two adjacent columns per lane (even nx): lane l owns
// columns 2l, 2l+1 of its 128-column strip, which is coarse column l, so the
// restriction's horizontal pair sum stays in the lane and the prolongation
// reads one coarse value per lane and row.  Tiles as the wide CG matvec
// (wide_tile, R even); the per-cell arithmetic and its order are those of
// the one-column kernels, so x, bc and z are identical to theirs.
template <class TB>
struct Pair;
template <>
struct Pair<float> {
    typedef f2v type;
};
template <>
struct Pair<double> {
    typedef d2v type;
};

template <class TB>
__global__ __launch_bounds__(256) void k_mg_down_march2(int nx, int ny, int nxc, int R, float omega,
                                                        const float* __restrict__ c, const float* __restrict__ e,
                                                        const float* __restrict__ n, const float* __restrict__ D,
                                                        const TB* __restrict__ b, float* __restrict__ x,
                                                        float* __restrict__ bc, const Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);  // iy0 even
    const int lane = threadIdx.x;
    const int ix = strip * 2 * WAVE + 2 * lane;  // columns ix, ix + 1
    const int iy1 = min(iy0 + R, ny);
    if (iy0 >= ny || strip * 2 * WAVE >= nx) return;  // uniform over the wave; no barriers follow
    const bool inx = ix < nx;
    const int ixc = inx ? ix : nx - 2;  // clamped pair for the loads of idle lanes
    auto xs2 = [&](size_t k) -> f2v {
        const TB2 bv = *(const TB2*)(b + k);
        const f2v dv = *(const f2v*)(D + k);
        return f2v{mg_div(omega * (float)bv.x, dv.x), mg_div(omega * (float)bv.y, dv.y)};
    };
    auto xs1 = [&](size_t k) -> float { return mg_div(omega * (float)b[k], D[k]); };
    f2v um = {0.f, 0.f}, nm = {0.f, 0.f};
    if (iy0 > 0) {
        um = xs2(cidx(ixc, iy0 - 1, nx));
        nm = *(const f2v*)(n + cidx(ixc, iy0 - 1, nx));
    }
    f2v uc = xs2(cidx(ixc, iy0, nx));
    float hrow = 0.f;  // the even row's pair sum
    for (int iy = iy0; iy < iy1; ++iy) {
        const size_t k = cidx(ixc, iy, nx);
        f2v up = {0.f, 0.f};
        if (iy + 1 < ny) up = xs2(k + nx);
        const f2v cc = *(const f2v*)(c + k), ec = *(const f2v*)(e + k), nc = *(const f2v*)(n + k);
        const TB2 bk = *(const TB2*)(b + k);
        float ue = lane_down1(uc.x);  // column ix + 2
        float uw = lane_up1(uc.y);    // column ix - 1
        float ew = lane_up1(ec.y);    // coupling (ix - 1, ix)
        if (lane == WAVE - 1 && ix + 2 < nx) ue = xs1(k + 2);
        if (lane == 0 && ix > 0) {
            uw = xs1(k - 1);
            ew = e[k - 1];
        }
        float a0 = cc.x * uc.x;
        a0 += ec.x * (uc.x - uc.y);  // ix < nx - 1 for every active pair
        if (ix > 0) a0 += ew * (uc.x - uw);
        if (iy < ny - 1) a0 += nc.x * (uc.x - up.x);
        if (iy > 0) a0 += nm.x * (uc.x - um.x);
        float a1 = cc.y * uc.y;
        if (ix + 1 < nx - 1) a1 += ec.y * (uc.y - ue);
        a1 += ec.x * (uc.y - uc.x);
        if (iy < ny - 1) a1 += nc.y * (uc.y - up.y);
        if (iy > 0) a1 += nm.y * (uc.y - um.y);
        const float r0 = inx ? (float)bk.x - a0 : 0.f, r1 = inx ? (float)bk.y - a1 : 0.f;
        if (inx) *(f2v*)(x + k) = uc;
        const float hs = r0 + r1;
        if ((iy & 1) == 0) {
            hrow = hs;
        } else if (inx) {
            bc[(size_t)(ix >> 1) + (size_t)(iy >> 1) * nxc] = hrow + hs;
        }
        um = uc;
        uc = up;
        nm = nc;
    }
    // odd ny: the last (even) row has no partner row
    if ((iy1 & 1) == 1 && inx) bc[(size_t)(ix >> 1) + (size_t)((iy1 - 1) >> 1) * nxc] = hrow + 0.f;
}

template <class TB>
__global__ __launch_bounds__(256) void k_mg_up_march2(int nx, int ny, int nxc, int R, float omega, float scale,
                                                      const float* __restrict__ c, const float* __restrict__ e,
                                                      const float* __restrict__ n, const float* __restrict__ D,
                                                      const TB* __restrict__ b, const float* __restrict__ x,
                                                      const float* __restrict__ xc, TB* __restrict__ out,
                                                      int with_dot, double* partials, Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);
    const int lane = threadIdx.x;
    const int ix = strip * 2 * WAVE + 2 * lane;
    const int iy1 = min(iy0 + R, ny);
    double acc = 0.0;
    if (iy0 < ny && strip * 2 * WAVE < nx) {  // uniform over the wave (the partial below needs every wave)
        const bool inx = ix < nx;
        const int ixc = inx ? ix : nx - 2;
        auto us2 = [&](int jx, int jy) -> f2v {
            const f2v xv = *(const f2v*)(x + (size_t)jx + (size_t)jy * nx);
            const float cv = xc[(size_t)(jx >> 1) + (size_t)(jy >> 1) * nxc];
            return f2v{xv.x + scale * cv, xv.y + scale * cv};
        };
        auto us1 = [&](int jx, int jy) -> float {
            return x[(size_t)jx + (size_t)jy * nx] + scale * xc[(size_t)(jx >> 1) + (size_t)(jy >> 1) * nxc];
        };
        f2v um = {0.f, 0.f}, nm = {0.f, 0.f};
        if (iy0 > 0) {
            um = us2(ixc, iy0 - 1);
            nm = *(const f2v*)(n + cidx(ixc, iy0 - 1, nx));
        }
        f2v uc = us2(ixc, iy0);
        for (int iy = iy0; iy < iy1; ++iy) {
            const size_t k = cidx(ixc, iy, nx);
            f2v up = {0.f, 0.f};
            if (iy + 1 < ny) up = us2(ixc, iy + 1);
            const f2v cc = *(const f2v*)(c + k), ec = *(const f2v*)(e + k), nc = *(const f2v*)(n + k),
                      dk = *(const f2v*)(D + k);
            const TB2 bk = *(const TB2*)(b + k);
            float ue = lane_down1(uc.x);
            float uw = lane_up1(uc.y);
            float ew = lane_up1(ec.y);
            if (lane == WAVE - 1 && ix + 2 < nx) ue = us1(ix + 2, iy);
            if (lane == 0 && ix > 0) {
                uw = us1(ix - 1, iy);
                ew = e[k - 1];
            }
            float a0 = cc.x * uc.x;
            a0 += ec.x * (uc.x - uc.y);
            if (ix > 0) a0 += ew * (uc.x - uw);
            if (iy < ny - 1) a0 += nc.x * (uc.x - up.x);
            if (iy > 0) a0 += nm.x * (uc.x - um.x);
            float a1 = cc.y * uc.y;
            if (ix + 1 < nx - 1) a1 += ec.y * (uc.y - ue);
            a1 += ec.x * (uc.y - uc.x);
            if (iy < ny - 1) a1 += nc.y * (uc.y - up.y);
            if (iy > 0) a1 += nm.y * (uc.y - um.y);
            const float z0 = uc.x + mg_div(omega * ((float)bk.x - a0), dk.x);
            const float z1 = uc.y + mg_div(omega * ((float)bk.y - a1), dk.y);
            if (inx) {
                TB2 zv;
                zv.x = (TB)z0;
                zv.y = (TB)z1;
                *(TB2*)(out + k) = zv;
                acc += (double)bk.x * (double)z0;
                acc += (double)bk.y * (double)z1;
            }
            um = uc;
            uc = up;
            nm = nc;
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// Coefficient pairs of the two-column red-black legs: FP32, or BF16 (the
// finest level, rounded to nearest even at set-up) unpacked from one 32-bit
// load; the element at the lower address is the low half.
// ldraw2 / unpack2 split the two, so a load can be issued a step before its
// value is unpacked.
__device__ __forceinline__ f2v ldraw2(const float* p, size_t k) { return *(const f2v*)(p + k); }
__device__ __forceinline__ uint32_t ldraw2(const uint16_t* p, size_t k) { return *(const uint32_t*)(p + k); }
__device__ __forceinline__ f2v unpack2(f2v v) { return v; }
__device__ __forceinline__ f2v unpack2(uint32_t w) {
    return f2v{__uint_as_float(w << 16), __uint_as_float(w & 0xFFFF0000u)};
}
__device__ __forceinline__ f2v ldc2(const float* p, size_t k) { return ldraw2(p, k); }
__device__ __forceinline__ f2v ldc2(const uint16_t* p, size_t k) { return unpack2(ldraw2(p, k)); }
// A coefficient pair as loaded: f2v for FP32, one 32-bit word for BF16.
template <class CT>
struct RawPair {
    typedef f2v type;
};
template <>
struct RawPair<uint16_t> {
    typedef uint32_t type;
};

// Red-black legs on two columns per lane (even nx).  A wave covers 128
// columns of which it owns the middle 124 (RB_OWN): lanes 0 and 63 are halo
// lanes that compute the neighbouring strips' two nearest columns by the
// same code, so every value an owned cell needs comes by shuffle (the halo
// lanes' own outer neighbours are unknown; nothing owned depends on them).
// Lane l holds columns ix, ix + 1 with ix = 124 strip - 2 + 2 l; in row j
// its red cell is column ix when j is even, ix + 1 when j is odd.
// Workgroups of four side-by-side strips over R rows (wide_tile), R even.
constexpr int RB_OWN = 2 * (WAVE - 2);

// Down-leg from x = 0: red, then black from its red neighbours; the wave runs
// two rows ahead (step t forms the red values of row t+2, the full x of row
// t+1 and the residual and restriction of row t, as k_mg_down_march2).  The
// loads of step t+1 (b, D of row t+3, aE, aN of row t+2, c of row t+1; row
// indices clamped into the grid, the values of rows outside it zeroed when
// unpacked) are issued before step t's arithmetic, so their latency overlaps
// it.
// UPD (FP64 CG with the deferred finish, TB = double): the CG update is
// fused in.  Every workgroup first takes k_cg_update_flat's prologue (the
// p'Ap and r'z partials, the zero-curvature test, alpha; workgroup 0 stores
// them), then the leg uses b - alpha Ap wherever it reads b and stores it to
// U.b_out for its own cells (a separate buffer: neighbouring tiles read the
// old b of the rows and columns they share).
template <class TB, class CT, bool UPD = false>
__global__ __launch_bounds__(256) void k_mg_down_rb2(int nx, int ny, int nxc, int R, float omega,
                                                     const CT* __restrict__ c, const CT* __restrict__ e,
                                                     const CT* __restrict__ n, const CT* __restrict__ D,
                                                     const TB* __restrict__ b, float* __restrict__ x,
                                                     float* __restrict__ bc, const Scalars* __restrict__ sc,
                                                     launch::CGUpd U) {
    typedef typename Pair<TB>::type TB2;
    typedef typename RawPair<CT>::type CR;
    static_assert(!UPD || sizeof(TB) == 8, "the fused update is FP64");
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);  // iy0 even
    const int lane = threadIdx.x;
    const int ix = strip * RB_OWN - 2 + 2 * lane;
    const int iy1 = min(iy0 + R, ny);
    const bool own = lane >= 1 && lane <= WAVE - 2 && ix < nx;
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);  // clamped pair for the loads of lanes outside
    const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;         // faces west of ix, east of ix + 1
    const f2v z2 = {0.f, 0.f};
    auto rowk = [&](int j) -> size_t { return cidx(ixc, j < 0 ? 0 : (j < ny ? j : ny - 1), nx); };
    // rows iy0-2 and iy0-1 (b, Ap, D) and n of row iy0-2: loaded (clamped into
    // the grid, so in bounds for every wave) before the update's prologue,
    // whose reductions they overlap; none of them needs alpha
    TB2 bv0[2];
    d2v av0[2];
    CR dv0[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const size_t k = rowk(iy0 - 2 + q);
        bv0[q] = *(const TB2*)(b + k);
        if constexpr (UPD) av0[q] = *(const d2v*)(U.Ap + k);
        dv0[q] = ldraw2(D, k);
    }
    const CR nv0 = ldraw2(n, rowk(iy0 - 2));
    // step t's loads: b, D of row t+2; aE, aN of row t+1; c of row t
    struct Ld {
        TB2 b;
        d2v ap;  // UPD: Ap of the same row
        CR d, e, n, c;
    };
    auto issue = [&](int t, Ld& q) {
        const size_t k2 = rowk(t + 2), k1 = rowk(t + 1);
        q.b = *(const TB2*)(b + k2);
        if constexpr (UPD) q.ap = *(const d2v*)(U.Ap + k2);
        q.d = ldraw2(D, k2);
        q.e = ldraw2(e, k1);
        q.n = ldraw2(n, k1);
        q.c = ldraw2(c, rowk(t));
    };
    // the first step's set goes out with them
    Ld qa, qb;
    issue(iy0 - 2, qa);
    double alpha = 0.0;
    if constexpr (UPD) {
        if (deferred_update_prologue(U.sc, U.pa, U.nba, U.pb, U.nbb, alpha)) return;
    }
    if (iy0 >= ny || strip * RB_OWN >= nx) return;  // uniform over the wave; no barriers follow
    // the red values of row iy0-2+q from its loaded b (with the fused update
    // applied, r - alpha Ap as the update computes it) and D
    auto red_row = [&](int q, f2v& bj, f2v& dj) -> f2v {
        const int j = iy0 - 2 + q;
        f2v r = z2;
        bj = z2;
        dj = z2;
        if (j < 0 || j >= ny) return r;
        TB2 bv = bv0[q];
        if constexpr (UPD) {
            bv.x = bv.x - alpha * av0[q].x;
            bv.y = bv.y - alpha * av0[q].y;
        }
        bj = f2v{(float)bv.x, (float)bv.y};  // b is used in FP32 only
        dj = unpack2(dv0[q]);
        if ((j & 1) == 0) r.x = mg_div(omega * bj.x, dj.x);
        else r.y = mg_div(omega * bj.y, dj.y);
        return r;
    };
    auto full_row = [&](int j, f2v rm, f2v rc, f2v rp, f2v bj, f2v dj, f2v ej, f2v nj, f2v nm) -> f2v {
        if (j < 0 || j >= ny) return z2;
        const bool hN = j < ny - 1, hS = j > 0;
        f2v v = rc;
        if ((j & 1) == 0) {  // black: column ix + 1
            const float xE = lane_down1(rc.x);
            v.y = rb_black0(bj.y, dj.y, omega, ej.x, rc.x, hE1 ? ej.y : 0.f, hE1 ? xE : 0.f, hN ? nj.y : 0.f,
                            hN ? rp.y : 0.f, hS ? nm.y : 0.f, hS ? rm.y : 0.f);
        } else {  // black: column ix
            const float xW = lane_up1(rc.y), aW = lane_up1(ej.y);
            v.x = rb_black0(bj.x, dj.x, omega, hW0 ? aW : 0.f, hW0 ? xW : 0.f, ej.x, rc.y, hN ? nj.x : 0.f,
                            hN ? rp.x : 0.f, hS ? nm.x : 0.f, hS ? rm.x : 0.f);
        }
        return v;
    };
    f2v b0, b1, d0, d1;
    f2v r0 = red_row(0, b0, d0);
    f2v r1 = red_row(1, b1, d1);
    f2v e0 = z2, n0 = z2, nm = z2, X0 = z2, Xm = z2;
    if (iy0 >= 2) n0 = unpack2(nv0);
    float hrow = 0.f;  // the even row's pair sum
    auto step = [&](int t, const Ld& q) {
        // red values of row t+2 (t+2 >= 0), as red_row
        f2v b2 = z2, d2 = z2, r2 = z2;
        if (t + 2 < ny) {
            TB2 bb = q.b;
            if constexpr (UPD) {
                bb.x = bb.x - alpha * q.ap.x;
                bb.y = bb.y - alpha * q.ap.y;
                if (own && t + 2 >= iy0 && t + 2 < iy1) *(TB2*)(U.b_out + cidx(ixc, t + 2, nx)) = bb;
            }
            b2 = f2v{(float)bb.x, (float)bb.y};
            d2 = unpack2(q.d);
            if (((t + 2) & 1) == 0) r2.x = mg_div(omega * b2.x, d2.x);
            else r2.y = mg_div(omega * b2.y, d2.y);
        }
        f2v e1 = z2, n1 = z2;
        if (t + 1 >= 0 && t + 1 < ny) {
            e1 = unpack2(q.e);
            n1 = unpack2(q.n);
        }
        const f2v X1 = full_row(t + 1, r0, r1, r2, b1, d1, e1, n1, n0);
        if (t >= iy0) {
            const size_t k = cidx(ixc, t, nx);
            const f2v cc = unpack2(q.c);
            const float ue = lane_down1(X0.x);  // column ix + 2
            const float uw = lane_up1(X0.y);    // column ix - 1
            const float ew = lane_up1(e0.y);    // coupling (ix - 1, ix)
            float a0 = cc.x * X0.x;
            a0 += e0.x * (X0.x - X0.y);  // ix < nx - 1 for every owned pair
            if (hW0) a0 += ew * (X0.x - uw);
            if (t < ny - 1) a0 += n0.x * (X0.x - X1.x);
            if (t > 0) a0 += nm.x * (X0.x - Xm.x);
            float a1 = cc.y * X0.y;
            if (hE1) a1 += e0.y * (X0.y - ue);
            a1 += e0.x * (X0.y - X0.x);
            if (t < ny - 1) a1 += n0.y * (X0.y - X1.y);
            if (t > 0) a1 += nm.y * (X0.y - Xm.y);
            const float q0 = own ? b0.x - a0 : 0.f, q1 = own ? b0.y - a1 : 0.f;
            if (own) *(f2v*)(x + k) = X0;
            const float hs = q0 + q1;
            if ((t & 1) == 0) {
                hrow = hs;
            } else if (own) {
                bc[(size_t)(ix >> 1) + (size_t)(t >> 1) * nxc] = hrow + hs;
            }
        }
        Xm = X0;
        X0 = X1;
        r0 = r1;
        r1 = r2;
        b0 = b1;
        b1 = b2;
        d1 = d2;
        nm = n0;
        n0 = n1;
        e0 = e1;
    };
    // two load sets used alternately, so that a set's loads stay in flight
    // across the loop's back edge (no register copies of loaded values)
    // (the last step's set loads rows that no step uses: clamped, in bounds,
    // issued unconditionally so that no branch around a load merges the wait
    // counts of the two paths)
    for (int t = iy0 - 2; t < iy1; t += 2) {
        issue(t + 1, qb);
        step(t, qa);
        if (t + 1 >= iy1) break;
        issue(t + 2, qa);
        step(t + 1, qb);
    }
    // odd ny: the last (even) row has no partner row
    if ((iy1 & 1) == 1 && own) bc[(size_t)(ix >> 1) + (size_t)((iy1 - 1) >> 1) * nxc] = hrow + 0.f;
}

// Up-leg: u = x + scale P xc; black cells relax from u, then red cells from
// the relaxed black neighbours.  Step t reads u of row t+2, relaxes the
// black cell of row t+1 and the red cell of row t, and writes row t.  (Loads
// issued a step ahead, as in the down-leg, measured slower here: 24.0 against
// 21.9 us at 4096 x 1024.)  Step
// t+1's loads are issued before step t's arithmetic, as in the down-leg.
template <class TB, class CT, class TO>
__global__ __launch_bounds__(256) void k_mg_up_rb2(int nx, int ny, int nxc, int R, float omega, float scale,
                                                   const CT* __restrict__ c, const CT* __restrict__ e,
                                                   const CT* __restrict__ n, const CT* __restrict__ D,
                                                   const TB* __restrict__ b, const float* __restrict__ x,
                                                   const float* __restrict__ xc, TO* __restrict__ out, int with_dot,
                                                   double* partials, Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);
    const int lane = threadIdx.x;
    const int ix = strip * RB_OWN - 2 + 2 * lane;
    const int iy1 = min(iy0 + R, ny);
    double acc = 0.0;
    if (iy0 < ny && strip * RB_OWN < nx) {  // uniform over the wave (the partial below needs every wave)
        const bool own = lane >= 1 && lane <= WAVE - 2 && ix < nx;
        const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
        const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;
        const f2v z2 = {0.f, 0.f};
        auto u_row = [&](int j) -> f2v {
            if (j < 0 || j >= ny) return z2;
            const f2v xv = *(const f2v*)(x + cidx(ixc, j, nx));
            const float cv = xc[(size_t)(ixc >> 1) + (size_t)(j >> 1) * nxc];
            return f2v{xv.x + scale * cv, xv.y + scale * cv};
        };
        auto black_row = [&](int j, f2v um, f2v uc, f2v up, f2v bj, f2v cj, f2v dj, f2v ej, f2v nj, f2v nm) -> f2v {
            if (j < 0 || j >= ny) return z2;
            const bool hN = j < ny - 1, hS = j > 0;
            f2v v = uc;
            if ((j & 1) == 0) {  // black: column ix + 1
                const float uE = lane_down1(uc.x);
                v.y = rb_relax(bj.y, cj.y, dj.y, omega, uc.y, hE1, hE1 ? ej.y : 0.f, hE1 ? uE : 0.f, true, ej.x,
                               uc.x, hN, hN ? nj.y : 0.f, hN ? up.y : 0.f, hS, hS ? nm.y : 0.f, hS ? um.y : 0.f);
            } else {  // black: column ix
                const float uW = lane_up1(uc.y), aW = lane_up1(ej.y);
                v.x = rb_relax(bj.x, cj.x, dj.x, omega, uc.x, true, ej.x, uc.y, hW0, hW0 ? aW : 0.f,
                               hW0 ? uW : 0.f, hN, hN ? nj.x : 0.f, hN ? up.x : 0.f, hS, hS ? nm.x : 0.f,
                               hS ? um.x : 0.f);
            }
            return v;
        };
        auto red_relax = [&](int j, f2v vm, f2v vc, f2v vp, f2v bj, f2v cj, f2v dj, f2v ej, f2v nj, f2v nm) -> f2v {
            const bool hN = j < ny - 1, hS = j > 0;
            f2v z = vc;
            if ((j & 1) == 0) {  // red: column ix
                const float vW = lane_up1(vc.y), aW = lane_up1(ej.y);
                z.x = rb_relax(bj.x, cj.x, dj.x, omega, vc.x, true, ej.x, vc.y, hW0, hW0 ? aW : 0.f,
                               hW0 ? vW : 0.f, hN, hN ? nj.x : 0.f, hN ? vp.x : 0.f, hS, hS ? nm.x : 0.f,
                               hS ? vm.x : 0.f);
            } else {  // red: column ix + 1
                const float vE = lane_down1(vc.x);
                z.y = rb_relax(bj.y, cj.y, dj.y, omega, vc.y, hE1, hE1 ? ej.y : 0.f, hE1 ? vE : 0.f, true, ej.x,
                               vc.x, hN, hN ? nj.y : 0.f, hN ? vp.y : 0.f, hS, hS ? nm.y : 0.f, hS ? vm.y : 0.f);
            }
            return z;
        };
        f2v u0 = u_row(iy0 - 2), u1 = u_row(iy0 - 1);
        f2v vm = z2, v0 = z2;
        TB2 b0;  // the FP64 (or FP32) rhs of row t, for the dot product
        b0.x = (TB)0;
        b0.y = (TB)0;
        f2v c0 = z2, d0 = z2, e0 = z2, n0 = z2, nm = z2;
        if (iy0 >= 2) n0 = ldc2(n, cidx(ixc, iy0 - 2, nx));
        for (int t = iy0 - 2; t < iy1; ++t) {
            const f2v u2 = u_row(t + 2);
            TB2 b1 = b0;
            f2v c1 = z2, d1 = z2, e1 = z2, n1 = z2;
            if (t + 1 >= 0 && t + 1 < ny) {
                const size_t k1 = cidx(ixc, t + 1, nx);
                b1 = *(const TB2*)(b + k1);
                c1 = ldc2(c, k1);
                d1 = ldc2(D, k1);
                e1 = ldc2(e, k1);
                n1 = ldc2(n, k1);
            }
            const f2v v1 = black_row(t + 1, u0, u1, u2, f2v{(float)b1.x, (float)b1.y}, c1, d1, e1, n1, n0);
            if (t >= iy0) {
                const f2v z = red_relax(t, vm, v0, v1, f2v{(float)b0.x, (float)b0.y}, c0, d0, e0, n0, nm);
                if (own) {
                    typename Pair<TO>::type zv;
                    zv.x = (TO)z.x;
                    zv.y = (TO)z.y;
                    *(typename Pair<TO>::type*)(out + cidx(ix, t, nx)) = zv;
                    acc += (double)b0.x * (double)z.x;
                    acc += (double)b0.y * (double)z.y;
                }
            }
            u0 = u1;
            u1 = u2;
            vm = v0;
            v0 = v1;
            b0 = b1;
            c0 = c1;
            d0 = d1;
            e0 = e1;
            nm = n0;
            n0 = n1;
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// The finest level's legs with two red-black sweeps each way (two columns per
// lane, marching; BF16 or FP32 coefficients; the fused CG update as in
// k_mg_down_rb2).  A marching step t takes five stages on five rows, each
// from the rows the stage before it left:
//   down: red from zero on row t+4, black from red on t+3, red relaxed on
//         t+2, black relaxed on t+1, the residual of row t;
//   up:   u = x + scale P xc on row t+4, black relaxed on t+3, red on t+2,
//         black on t+1, red on t (z).
// Each stage changes one colour of its row and reads the other colour of the
// rows beside it, so one register window of the current x serves all five.
// A value formed by a half-sweep is wrong one column further in from the
// wave's edge than what it reads: four half-sweeps and the residual need two
// halo lanes a side (120 owned columns of the wave's 128).  Every cell is
// formed with the arithmetic of the one-sweep legs in the same order; later
// sweeps relax with rb_relax.
constexpr int SW2_HALO = 2, SW2_OWN = 2 * (WAVE - 2 * SW2_HALO);
// a marching step's phase in a loop unrolled six times, as a type
template <int P>
struct Phase {
    static constexpr int value = P;
};

// red (or black) cells of a row pair relaxed in place from the rows beside it.
// Every east and west term is taken: a coupling across the grid's edge is
// exactly zero (the assembly writes aE = 0 on the last column, and the legs
// zero the couplings of the lanes west of column 0), so it adds zero times a
// finite value, and no lane-dependent select is needed.
__device__ __forceinline__ void sw2_relax(f2v& uc, f2v um, f2v up, int j, bool red_pass, int ny, float omega,
                                          f2v bj, f2v cj, f2v dj, f2v ej, f2v nj, f2v nm) {
    const bool hN = j < ny - 1, hS = j > 0;
    if (((j & 1) == 0) == red_pass) {
        const float uW = lane_up1(uc.y), aW = lane_up1(ej.y);
        uc.x = rb_relax(bj.x, cj.x, dj.x, omega, uc.x, true, ej.x, uc.y, true, aW, uW, hN, hN ? nj.x : 0.f,
                        hN ? up.x : 0.f, hS, hS ? nm.x : 0.f, hS ? um.x : 0.f);
    } else {
        const float uE = lane_down1(uc.x);
        uc.y = rb_relax(bj.y, cj.y, dj.y, omega, uc.y, true, ej.y, uE, true, ej.x, uc.x, hN, hN ? nj.y : 0.f,
                        hN ? up.y : 0.f, hS, hS ? nm.y : 0.f, hS ? um.y : 0.f);
    }
}

template <class TB, class CT, bool UPD = false>
__global__ __launch_bounds__(256) void k_mg_down_sw2(int nx, int ny, int nxc, int R, float omega,
                                                     const CT* __restrict__ c, const CT* __restrict__ e,
                                                     const CT* __restrict__ n, const CT* __restrict__ D,
                                                     const TB* __restrict__ b, float* __restrict__ x,
                                                     float* __restrict__ bc, const Scalars* __restrict__ sc,
                                                     launch::CGUpd U) {
    typedef typename Pair<TB>::type TB2;
    typedef typename RawPair<CT>::type CR;
    static_assert(!UPD || sizeof(TB) == 8, "the fused update is FP64");
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);  // iy0 even
    const int lane = threadIdx.x;
    const int ix = strip * SW2_OWN - 2 * SW2_HALO + 2 * lane;
    const int iy1 = min(iy0 + R, ny);
    const bool own = lane >= SW2_HALO && lane <= WAVE - 1 - SW2_HALO && ix < nx;
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);  // clamped pair for the loads of lanes outside
    const f2v z2 = {0.f, 0.f};
    auto rowk = [&](int j) -> size_t { return cidx(ixc, j < 0 ? 0 : (j < ny ? j : ny - 1), nx); };
    auto in = [&](int j) { return j >= 0 && j < ny; };
    // step t's loads: b (with Ap) and D of row t+4, e and n of row t+3, c of row t+2
    struct Ld {
        TB2 b;
        d2v ap;  // UPD: Ap of the same row
        CR d, e, n, c;
    };
    auto issue = [&](int t, Ld& q) {
        const size_t k4 = rowk(t + 4), k3 = rowk(t + 3);
        q.b = *(const TB2*)(b + k4);
        if constexpr (UPD) q.ap = *(const d2v*)(U.Ap + k4);
        q.d = ldraw2(D, k4);
        q.e = ldraw2(e, k3);
        q.n = ldraw2(n, k3);
        q.c = ldraw2(c, rowk(t + 2));
    };
    const int t0 = iy0 - 8;  // the first row whose red cells row iy0's residual depends on is iy0 - 4
    Ld qa, qb;
    issue(t0, qa);  // in flight across the update's prologue
    double alpha = 0.0;
    if constexpr (UPD) {
        if (deferred_update_prologue(U.sc, U.pa, U.nba, U.pb, U.nbb, alpha)) return;
    }
    if (iy0 >= ny || strip * SW2_OWN >= nx) return;  // uniform over the wave; no barriers follow
    // six-row ring windows (x and every coefficient): at marching step t, row
    // t + o (o = -1 .. 4) sits in slot (P + 1 + o) % 6, P = (t - t0) % 6, so a
    // step names its rows by fixed slots and nothing moves between steps; the
    // loop below is unrolled six times.  A row enters at o = 4 (b, D, x), 3
    // (e, n) or 2 (c) and is last read at o = -1 (x, n) or o = 0.
    f2v X[6], Bw[6], Dw[6], Ew[6], Nw[6], Cw[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) X[k] = Bw[k] = Dw[k] = Ew[k] = Nw[k] = Cw[k] = z2;
    float hrow = 0.f;  // the even row's pair sum
    auto step = [&](auto ph, int t, const Ld& q) {
        constexpr int P = decltype(ph)::value;
        constexpr int om = P % 6, o0 = (P + 1) % 6, o1 = (P + 2) % 6, o2 = (P + 3) % 6, o3 = (P + 4) % 6,
                      o4 = (P + 5) % 6;
        const int j4 = t + 4, j3 = t + 3, j2 = t + 2, j1 = t + 1;
        // row t+4: b (the fused update applied) and D, and its red cells from zero
        f2v b4 = z2, d4 = z2, x4 = z2;
        if (in(j4)) {
            TB2 bb = q.b;
            if constexpr (UPD) {
                bb.x = bb.x - alpha * q.ap.x;
                bb.y = bb.y - alpha * q.ap.y;
                if (own && j4 >= iy0 && j4 < iy1) *(TB2*)(U.b_out + cidx(ixc, j4, nx)) = bb;
            }
            b4 = f2v{(float)bb.x, (float)bb.y};  // b is used in FP32 only
            d4 = unpack2(q.d);
            if ((j4 & 1) == 0) x4.x = mg_div(omega * b4.x, d4.x);
            else x4.y = mg_div(omega * b4.y, d4.y);
        }
        X[o4] = x4;
        Bw[o4] = b4;
        Dw[o4] = d4;
        Ew[o3] = in(j3) && ix >= 0 ? unpack2(q.e) : z2;  // lanes west of column 0: no coupling
        Nw[o3] = in(j3) ? unpack2(q.n) : z2;
        Cw[o2] = in(j2) ? unpack2(q.c) : z2;
        // row t+3: black cells from the red ones
        if (in(j3)) {
            const f2v rm = X[o2], rc = X[o3], rp = X[o4], bj = Bw[o3], dj = Dw[o3], ej = Ew[o3], nj = Nw[o3],
                      nm = Nw[o2];
            const bool hN = j3 < ny - 1, hS = j3 > 0;
            if ((j3 & 1) == 0) {
                const float xE = lane_down1(rc.x);
                X[o3].y = rb_black0(bj.y, dj.y, omega, ej.x, rc.x, ej.y, xE, hN ? nj.y : 0.f, hN ? rp.y : 0.f,
                                    hS ? nm.y : 0.f, hS ? rm.y : 0.f);
            } else {
                const float xW = lane_up1(rc.y), aW = lane_up1(ej.y);
                X[o3].x = rb_black0(bj.x, dj.x, omega, aW, xW, ej.x, rc.y, hN ? nj.x : 0.f, hN ? rp.x : 0.f,
                                    hS ? nm.x : 0.f, hS ? rm.x : 0.f);
            }
        }
        // row t+2: red cells relaxed; row t+1: black cells relaxed
        if (in(j2))
            sw2_relax(X[o2], X[o1], X[o3], j2, true, ny, omega, Bw[o2], Cw[o2], Dw[o2], Ew[o2], Nw[o2],
                      Nw[o1]);
        if (in(j1))
            sw2_relax(X[o1], X[o0], X[o2], j1, false, ny, omega, Bw[o1], Cw[o1], Dw[o1], Ew[o1], Nw[o1],
                      Nw[o0]);
        // row t: the residual of the final x, restricted by 2 x 2 sums
        if (t >= iy0) {
            const size_t k = cidx(ixc, t, nx);
            const f2v X0 = X[o0], Xm = X[om], X1 = X[o1], cc = Cw[o0], e0 = Ew[o0], n0 = Nw[o0], nm = Nw[om],
                      b0 = Bw[o0];
            const float ue = lane_down1(X0.x);  // column ix + 2
            const float uw = lane_up1(X0.y);    // column ix - 1
            const float ew = lane_up1(e0.y);    // coupling (ix - 1, ix)
            float a0 = cc.x * X0.x;
            a0 += e0.x * (X0.x - X0.y);
            a0 += ew * (X0.x - uw);
            if (t < ny - 1) a0 += n0.x * (X0.x - X1.x);
            if (t > 0) a0 += nm.x * (X0.x - Xm.x);
            float a1 = cc.y * X0.y;
            a1 += e0.y * (X0.y - ue);
            a1 += e0.x * (X0.y - X0.x);
            if (t < ny - 1) a1 += n0.y * (X0.y - X1.y);
            if (t > 0) a1 += nm.y * (X0.y - Xm.y);
            const float q0 = own ? b0.x - a0 : 0.f, q1 = own ? b0.y - a1 : 0.f;
            if (own) *(f2v*)(x + k) = X0;
            const float hs = q0 + q1;
            if ((t & 1) == 0) {
                hrow = hs;
            } else if (own) {
                bc[(size_t)(ix >> 1) + (size_t)(t >> 1) * nxc] = hrow + hs;
            }
        }
    };
    // two load sets used alternately, so that a set's loads stay in flight
    // across the step before it (the last step's set loads rows that no step
    // uses: clamped, in bounds)
    for (int t = t0; t < iy1; t += 6) {
        issue(t + 1, qb);
        step(Phase<0>{}, t, qa);
        if (t + 1 >= iy1) break;
        issue(t + 2, qa);
        step(Phase<1>{}, t + 1, qb);
        if (t + 2 >= iy1) break;
        issue(t + 3, qb);
        step(Phase<2>{}, t + 2, qa);
        if (t + 3 >= iy1) break;
        issue(t + 4, qa);
        step(Phase<3>{}, t + 3, qb);
        if (t + 4 >= iy1) break;
        issue(t + 5, qb);
        step(Phase<4>{}, t + 4, qa);
        if (t + 5 >= iy1) break;
        issue(t + 6, qa);
        step(Phase<5>{}, t + 5, qb);
    }
    // odd ny: the last (even) row has no partner row
    if ((iy1 & 1) == 1 && own) bc[(size_t)(ix >> 1) + (size_t)((iy1 - 1) >> 1) * nxc] = hrow + 0.f;
}

template <class TB, class CT, class TO>
__global__ __launch_bounds__(256) void k_mg_up_sw2(int nx, int ny, int nxc, int R, float omega, float scale,
                                                   const CT* __restrict__ c, const CT* __restrict__ e,
                                                   const CT* __restrict__ n, const CT* __restrict__ D,
                                                   const TB* __restrict__ b, const float* __restrict__ x,
                                                   const float* __restrict__ xc, TO* __restrict__ out, int with_dot,
                                                   double* partials, Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    if (sc->done) return;
    int strip, iy0;
    wide_tile(R, strip, iy0);
    const int lane = threadIdx.x;
    const int ix = strip * SW2_OWN - 2 * SW2_HALO + 2 * lane;
    const int iy1 = min(iy0 + R, ny);
    double acc = 0.0;
    if (iy0 < ny && strip * SW2_OWN < nx) {  // uniform over the wave (the partial below needs every wave)
        const bool own = lane >= SW2_HALO && lane <= WAVE - 1 - SW2_HALO && ix < nx;
        const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
        const f2v z2 = {0.f, 0.f};
        auto in = [&](int j) { return j >= 0 && j < ny; };
        // six-row ring windows as in k_mg_down_sw2: row t + o of step t in slot
        // (P + 1 + o) % 6, P = (t - t0) % 6; u enters at o = 4, b and the
        // coefficients at o = 3
        f2v Uw[6], Bw[6], Cw[6], Dw[6], Ew[6], Nw[6];
        TB2 Braw[6];  // b in its own precision, for the dot product
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            Uw[k] = Bw[k] = Cw[k] = Dw[k] = Ew[k] = Nw[k] = z2;
            Braw[k].x = (TB)0;
            Braw[k].y = (TB)0;
        }
        // step t's loads: x and xc of row t+4; b, c, D, e, n of row t+3 (rows
        // clamped into the grid; rows outside it are zeroed where they are used)
        typedef typename RawPair<CT>::type CR;
        struct Ld {
            f2v x;
            float xc;
            TB2 b;
            CR c, d, e, n;
        };
        auto rowk = [&](int j) -> int { return j < 0 ? 0 : (j < ny ? j : ny - 1); };
        auto issue = [&](int t, Ld& q) {
            const int r4 = rowk(t + 4);
            const size_t k3 = cidx(ixc, rowk(t + 3), nx);
            q.x = *(const f2v*)(x + cidx(ixc, r4, nx));
            q.xc = xc[(size_t)(ixc >> 1) + (size_t)(r4 >> 1) * nxc];
            q.b = *(const TB2*)(b + k3);
            q.c = ldraw2(c, k3);
            q.d = ldraw2(D, k3);
            q.e = ldraw2(e, k3);
            q.n = ldraw2(n, k3);
        };
        auto step = [&](auto ph, int t, const Ld& q) {
            constexpr int P = decltype(ph)::value;
            constexpr int om = P % 6, o0 = (P + 1) % 6, o1 = (P + 2) % 6, o2 = (P + 3) % 6, o3 = (P + 4) % 6,
                          o4 = (P + 5) % 6;
            const int j4 = t + 4, j3 = t + 3, j2 = t + 2, j1 = t + 1;
            Uw[o4] = in(j4) ? f2v{q.x.x + scale * q.xc, q.x.y + scale * q.xc} : z2;
            TB2 b3;
            b3.x = (TB)0;
            b3.y = (TB)0;
            f2v c3 = z2, d3 = z2, e3 = z2, n3 = z2;
            if (in(j3)) {
                b3 = q.b;
                c3 = unpack2(q.c);
                d3 = unpack2(q.d);
                e3 = ix >= 0 ? unpack2(q.e) : z2;  // lanes west of column 0: no coupling
                n3 = unpack2(q.n);
            }
            Braw[o3] = b3;
            Bw[o3] = f2v{(float)b3.x, (float)b3.y};
            Cw[o3] = c3;
            Dw[o3] = d3;
            Ew[o3] = e3;
            Nw[o3] = n3;
            // black on t+3 (from u), red on t+2, black on t+1, red on t
            if (in(j3))
                sw2_relax(Uw[o3], Uw[o2], Uw[o4], j3, false, ny, omega, Bw[o3], Cw[o3], Dw[o3], Ew[o3],
                          Nw[o3], Nw[o2]);
            if (in(j2))
                sw2_relax(Uw[o2], Uw[o1], Uw[o3], j2, true, ny, omega, Bw[o2], Cw[o2], Dw[o2], Ew[o2],
                          Nw[o2], Nw[o1]);
            if (in(j1))
                sw2_relax(Uw[o1], Uw[o0], Uw[o2], j1, false, ny, omega, Bw[o1], Cw[o1], Dw[o1], Ew[o1],
                          Nw[o1], Nw[o0]);
            if (t >= iy0) {
                sw2_relax(Uw[o0], Uw[om], Uw[o1], t, true, ny, omega, Bw[o0], Cw[o0], Dw[o0], Ew[o0],
                          Nw[o0], Nw[om]);
                if (own) {
                    const f2v z = Uw[o0];
                    typename Pair<TO>::type zv;
                    zv.x = (TO)z.x;
                    zv.y = (TO)z.y;
                    *(typename Pair<TO>::type*)(out + cidx(ix, t, nx)) = zv;
                    acc += (double)Braw[o0].x * (double)z.x;
                    acc += (double)Braw[o0].y * (double)z.y;
                }
            }
        };
        // two load sets used alternately, a step ahead (as in the down-leg)
        const int t0 = iy0 - 8;
        Ld qa, qb;
        issue(t0, qa);
        for (int t = t0; t < iy1; t += 6) {
            issue(t + 1, qb);
            step(Phase<0>{}, t, qa);
            if (t + 1 >= iy1) break;
            issue(t + 2, qa);
            step(Phase<1>{}, t + 1, qb);
            if (t + 2 >= iy1) break;
            issue(t + 3, qb);
            step(Phase<2>{}, t + 2, qa);
            if (t + 3 >= iy1) break;
            issue(t + 4, qa);
            step(Phase<3>{}, t + 3, qb);
            if (t + 4 >= iy1) break;
            issue(t + 5, qb);
            step(Phase<4>{}, t + 4, qa);
            if (t + 5 >= iy1) break;
            issue(t + 6, qa);
            step(Phase<5>{}, t + 5, qb);
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// The same two legs for the levels where a wave takes two rows (R = 2):
// these levels are latency-bound, so every load of the wave's rows (row
// indices clamped into the grid, rows outside it zeroed after the load) is
// issued at once, and the done flag is read with them; the arithmetic per
// row is that of k_mg_down_rb2 / k_mg_up_rb2, in the same order.
// LIN: the level reaches the next one by the linear transfers.  The
// restriction of coarse row J = iy0 / 2 gathers the residuals of fine rows
// iy0-1 .. iy0+2 and, in each, of the columns one beyond the lane's pair on
// either side; so the wave forms one more row of x above and below (E = 1)
// and keeps two halo lanes on each side of its owned columns (RB_OWN_LIN):
// the inner halo lane's x is exact (its black cell's outer red neighbour is
// in the outer halo lane), and so is its residual.
constexpr int RB_OWN_LIN = 2 * (WAVE - 4);
template <bool LIN>
struct RBLanes {
    static constexpr int HALO = LIN ? 2 : 1, OWN = 2 * (WAVE - 2 * HALO);
};
template <class TB, class CT, bool LIN = false>
__global__ __launch_bounds__(256) void k_mg_down_rb2s(int nx, int ny, int nxc, float omega,
                                                      const CT* __restrict__ c, const CT* __restrict__ e,
                                                      const CT* __restrict__ n, const CT* __restrict__ D,
                                                      const TB* __restrict__ b, float* __restrict__ x,
                                                      float* __restrict__ bc, const Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    constexpr int E = LIN ? 1 : 0, HALO = RBLanes<LIN>::HALO, OWN = RBLanes<LIN>::OWN;
    constexpr int NB = 6 + 2 * E, NN = 5 + 2 * E, NE = 4 + 2 * E, NC = 2 + 2 * E, NX = 4 + 2 * E;
    const int done = sc->done;  // arrives with the loads below
    int strip, iy0;
    wide_tile(2, strip, iy0);  // iy0 even
    const int lane = threadIdx.x;
    const int ix = strip * OWN - 2 * HALO + 2 * lane;
    if (iy0 >= ny || strip * OWN >= nx) return;  // uniform over the wave; no barriers follow
    const bool own = lane >= HALO && lane <= WAVE - 1 - HALO && ix < nx;
    const bool inx = ix >= 0 && ix < nx;
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
    const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;
    const f2v z2 = {0.f, 0.f};
    auto row = [&](int j) -> size_t { return cidx(ixc, min(max(j, 0), ny - 1), nx); };
    // b and D of rows iy0-2-E .. iy0+3+E, n of rows iy0-2-E .. iy0+2+E, e of
    // rows iy0-1-E .. iy0+2+E, c of rows iy0-E .. iy0+1+E
    const int jb = iy0 - 2 - E;
    f2v Bv[NB], Dv[NB], Nv[NN], Ev[NE], Cv[NC];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const size_t k = row(jb + q);
        const TB2 v = *(const TB2*)(b + k);
        Bv[q] = f2v{(float)v.x, (float)v.y};
        Dv[q] = ldc2(D, k);
    }
#pragma unroll
    for (int q = 0; q < NN; ++q) Nv[q] = ldc2(n, row(jb + q));
#pragma unroll
    for (int q = 0; q < NE; ++q) Ev[q] = ldc2(e, row(jb + 1 + q));
#pragma unroll
    for (int q = 0; q < NC; ++q) Cv[q] = ldc2(c, row(jb + 2 + q));
    // every load is issued before the flag decides
#pragma unroll
    for (int q = 0; q < NB; ++q) asm volatile("" : : "v"(Bv[q].x), "v"(Bv[q].y), "v"(Dv[q].x), "v"(Dv[q].y));
#pragma unroll
    for (int q = 0; q < NN; ++q) asm volatile("" : : "v"(Nv[q].x), "v"(Nv[q].y));
#pragma unroll
    for (int q = 0; q < NE; ++q) asm volatile("" : : "v"(Ev[q].x), "v"(Ev[q].y));
#pragma unroll
    for (int q = 0; q < NC; ++q) asm volatile("" : : "v"(Cv[q].x), "v"(Cv[q].y));
    if (done) return;
    f2v Rv[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int j = jb + q;
        Rv[q] = z2;
        if (j >= 0 && j < ny) {
            if ((j & 1) == 0) Rv[q].x = mg_div(omega * Bv[q].x, Dv[q].x);
            else Rv[q].y = mg_div(omega * Bv[q].y, Dv[q].y);
        }
        if (j < 0 || j >= ny) {
            Bv[q] = z2;
            Dv[q] = z2;
        }
    }
    // full x of rows iy0-1-E .. iy0+2+E
    f2v Xv[NX];
#pragma unroll
    for (int q = 1; q <= NX; ++q) {
        const int j = jb + q;
        f2v v = z2;
        if (j >= 0 && j < ny) {
            const f2v rm = Rv[q - 1], rc = Rv[q], rp = Rv[q + 1], ej = Ev[q - 1], nj = Nv[q], nm = Nv[q - 1];
            const bool hN = j < ny - 1, hS = j > 0;
            v = rc;
            if ((j & 1) == 0) {
                const float xE = lane_down1(rc.x);
                v.y = rb_black0(Bv[q].y, Dv[q].y, omega, ej.x, rc.x, hE1 ? ej.y : 0.f, hE1 ? xE : 0.f,
                                hN ? nj.y : 0.f, hN ? rp.y : 0.f, hS ? nm.y : 0.f, hS ? rm.y : 0.f);
            } else {
                const float xW = lane_up1(rc.y), aW = lane_up1(ej.y);
                v.x = rb_black0(Bv[q].x, Dv[q].x, omega, hW0 ? aW : 0.f, hW0 ? xW : 0.f, ej.x, rc.y,
                                hN ? nj.x : 0.f, hN ? rp.x : 0.f, hS ? nm.x : 0.f, hS ? rm.x : 0.f);
            }
        }
        Xv[q - 1] = v;
    }
    // the residual of rows iy0-E .. iy0+1+E (row t: X row index t - jb - 1)
    float hrow = 0.f, acc = 0.f;
    const float wx0 = ix == 0 ? 1.f : 0.75f, wx1 = ix + 1 == nx - 1 ? 1.f : 0.75f;
#pragma unroll
    for (int s2 = 0; s2 < 2 + 2 * E; ++s2) {
        const int t = iy0 - E + s2;
        if (t >= 0 && t < ny) {  // uniform over the wave
            const f2v X0 = Xv[s2 + 1], Xm = Xv[s2], X1 = Xv[s2 + 2];
            const f2v e0 = Ev[s2 + 1], n0 = Nv[s2 + 2], nm = Nv[s2 + 1], cc = Cv[s2], b0 = Bv[s2 + 2];
            const float ue = lane_down1(X0.x);
            const float uw = lane_up1(X0.y);
            const float ew = lane_up1(e0.y);
            float a0 = cc.x * X0.x;
            a0 += e0.x * (X0.x - X0.y);
            if (hW0) a0 += ew * (X0.x - uw);
            if (t < ny - 1) a0 += n0.x * (X0.x - X1.x);
            if (t > 0) a0 += nm.x * (X0.x - Xm.x);
            float a1 = cc.y * X0.y;
            if (hE1) a1 += e0.y * (X0.y - ue);
            a1 += e0.x * (X0.y - X0.x);
            if (t < ny - 1) a1 += n0.y * (X0.y - X1.y);
            if (t > 0) a1 += nm.y * (X0.y - Xm.y);
            if constexpr (LIN) {
                // P^T: this row's residuals restricted along x (columns
                // ix-1 .. ix+2), weighted into coarse row iy0 / 2
                const float q0 = inx ? b0.x - a0 : 0.f, q1 = inx ? b0.y - a1 : 0.f;
                const float rl = lane_up1(q1), rr = lane_down1(q0);
                float h = 0.25f * rl;
                h += wx0 * q0;
                h += wx1 * q1;
                h += 0.25f * rr;
                acc += lin_rw(t, s2, ny) * h;
                if (own && (t == iy0 || t == iy0 + 1)) *(f2v*)(x + cidx(ix, t, nx)) = X0;
            } else {
                const float q0 = own ? b0.x - a0 : 0.f, q1 = own ? b0.y - a1 : 0.f;
                if (own) *(f2v*)(x + cidx(ix, t, nx)) = X0;
                const float hs = q0 + q1;
                if (s2 == 0) hrow = hs;
                else if (own) bc[(size_t)(ix >> 1) + (size_t)(t >> 1) * nxc] = hrow + hs;
            }
        }
    }
    if constexpr (LIN) {
        if (own) bc[(size_t)(ix >> 1) + (size_t)(iy0 >> 1) * nxc] = acc;
    } else {
        // odd ny: the last (even) row has no partner row
        if (iy0 + 1 >= ny && own) bc[(size_t)(ix >> 1) + (size_t)(iy0 >> 1) * nxc] = hrow + 0.f;
    }
}

// LIN: u = x + scale P xc with the linear prolongation: each lane reads its
// coarse column in coarse rows iy0/2-2 .. iy0/2+2, interpolates along y, and
// takes its neighbours' values along x by shuffle (two halo lanes a side, as
// the down-leg: the outer halo lane's outer column is wrong and nothing owned
// depends on it).
template <class TB, class CT, class TO, bool LIN = false>
__global__ __launch_bounds__(256) void k_mg_up_rb2s(int nx, int ny, int nxc, float omega, float scale,
                                                    const CT* __restrict__ c, const CT* __restrict__ e,
                                                    const CT* __restrict__ n, const CT* __restrict__ D,
                                                    const TB* __restrict__ b, const float* __restrict__ x,
                                                    const float* __restrict__ xc, TO* __restrict__ out, int with_dot,
                                                    double* partials, Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    constexpr int HALO = RBLanes<LIN>::HALO, OWN = RBLanes<LIN>::OWN;
    const int done = sc->done;  // arrives with the loads below
    int strip, iy0;
    wide_tile(2, strip, iy0);
    const int lane = threadIdx.x;
    const int ix = strip * OWN - 2 * HALO + 2 * lane;
    double acc = 0.0;
    const bool active = iy0 < ny && strip * OWN < nx;  // uniform over the wave
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
    const int iyc = active ? iy0 : 0;  // rows of an inactive wave: any in-grid rows (never used)
    auto row = [&](int j) -> size_t { return cidx(ixc, min(max(j, 0), ny - 1), nx); };
    // u (x and xc) of rows iy0-2 .. iy0+3, b c D e of rows iy0-1 .. iy0+2, n of rows iy0-2 .. iy0+2
    // (LIN: xc of coarse rows iy0/2-2 .. iy0/2+2 in Lc)
    f2v Xv[6], Cx[6];
    float Lc[5];
    TB2 Bt[4];
    f2v Cv[4], Dv[4], Ev[4], Nv[5];
    const int nyc = (ny + 1) >> 1;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int j = min(max(iyc - 2 + q, 0), ny - 1);
        Xv[q] = *(const f2v*)(x + cidx(ixc, j, nx));
        if constexpr (!LIN) {
            const float cv = xc[(size_t)(ixc >> 1) + (size_t)(j >> 1) * nxc];
            Cx[q] = f2v{cv, cv};
        }
    }
    if constexpr (LIN) {
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int J = min(max((iyc >> 1) - 2 + q, 0), nyc - 1);
            Lc[q] = xc[(size_t)(ixc >> 1) + (size_t)J * nxc];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const size_t k = row(iyc - 1 + q);
        Bt[q] = *(const TB2*)(b + k);
        Cv[q] = ldc2(c, k);
        Dv[q] = ldc2(D, k);
        Ev[q] = ldc2(e, k);
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) Nv[q] = ldc2(n, row(iyc - 2 + q));
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        asm volatile("" : : "v"(Xv[q].x), "v"(Xv[q].y));
        if constexpr (!LIN) asm volatile("" : : "v"(Cx[q].x));
    }
    if constexpr (LIN) {
#pragma unroll
        for (int q = 0; q < 5; ++q) asm volatile("" : : "v"(Lc[q]));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        asm volatile("" : : "v"(Bt[q].x), "v"(Bt[q].y), "v"(Cv[q].x), "v"(Cv[q].y), "v"(Dv[q].x), "v"(Dv[q].y),
                     "v"(Ev[q].x), "v"(Ev[q].y));
#pragma unroll
    for (int q = 0; q < 5; ++q) asm volatile("" : : "v"(Nv[q].x), "v"(Nv[q].y));
    if (done) return;  // uniform over the workgroup, before the partial's barriers
    if (active) {
        const bool own = lane >= HALO && lane <= WAVE - 1 - HALO && ix < nx;
        const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;
        const f2v z2 = {0.f, 0.f};
        f2v U[6];
        if constexpr (LIN) {
            // fine row j = iy0-2+q: own coarse row J0 = iy0/2-1+q/2 (Lc index
            // 1 + q/2), neighbour J0 - 1 (even j) or J0 + 1 (odd j), folded
            const float wl = ix == 0 ? 0.f : 0.25f, wr = ix + 1 == nx - 1 ? 0.f : 0.25f;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int j = iy0 - 2 + q;
                const int i0 = 1 + (q >> 1), i1 = (q & 1) ? i0 + 1 : i0 - 1;
                const int J1 = (iy0 >> 1) - 2 + i1;
                const bool fold = J1 < 0 || J1 > nyc - 1;
                const float v = fold ? Lc[i0] : 0.75f * Lc[i0] + 0.25f * Lc[i1];
                const float vl = lane_up1(v), vr = lane_down1(v);
                const float p0 = (1.f - wl) * v + wl * vl, p1 = (1.f - wr) * v + wr * vr;
                U[q] = (j >= 0 && j < ny) ? f2v{Xv[q].x + scale * p0, Xv[q].y + scale * p1} : z2;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int j = iy0 - 2 + q;
                U[q] = (j >= 0 && j < ny) ? f2v{Xv[q].x + scale * Cx[q].x, Xv[q].y + scale * Cx[q].y} : z2;
            }
        }
        // rows iy0-1 .. iy0+2 with their black cell relaxed
        f2v V[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = iy0 - 1 + q;
            f2v v = z2;
            if (j >= 0 && j < ny) {
                const f2v um = U[q], uc = U[q + 1], up = U[q + 2], cj = Cv[q], dj = Dv[q], ej = Ev[q];
                const f2v nj = Nv[q + 1], nm = Nv[q];
                const f2v bj = f2v{(float)Bt[q].x, (float)Bt[q].y};
                const bool hN = j < ny - 1, hS = j > 0;
                v = uc;
                if ((j & 1) == 0) {
                    const float uE = lane_down1(uc.x);
                    v.y = rb_relax(bj.y, cj.y, dj.y, omega, uc.y, hE1, hE1 ? ej.y : 0.f, hE1 ? uE : 0.f, true,
                                   ej.x, uc.x, hN, hN ? nj.y : 0.f, hN ? up.y : 0.f, hS, hS ? nm.y : 0.f,
                                   hS ? um.y : 0.f);
                } else {
                    const float uW = lane_up1(uc.y), aW = lane_up1(ej.y);
                    v.x = rb_relax(bj.x, cj.x, dj.x, omega, uc.x, true, ej.x, uc.y, hW0, hW0 ? aW : 0.f,
                                   hW0 ? uW : 0.f, hN, hN ? nj.x : 0.f, hN ? up.x : 0.f, hS, hS ? nm.x : 0.f,
                                   hS ? um.x : 0.f);
                }
            }
            V[q] = v;
        }
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            const int t = iy0 + s2;
            if (t < ny) {
                const f2v vm = V[s2], vc = V[s2 + 1], vp = V[s2 + 2];
                const f2v cj = Cv[s2 + 1], dj = Dv[s2 + 1], ej = Ev[s2 + 1], nj = Nv[s2 + 2], nm = Nv[s2 + 1];
                const TB2 bt = Bt[s2 + 1];
                const f2v bj = f2v{(float)bt.x, (float)bt.y};
                const bool hN = t < ny - 1, hS = t > 0;
                f2v z = vc;
                if ((t & 1) == 0) {
                    const float vW = lane_up1(vc.y), aW = lane_up1(ej.y);
                    z.x = rb_relax(bj.x, cj.x, dj.x, omega, vc.x, true, ej.x, vc.y, hW0, hW0 ? aW : 0.f,
                                   hW0 ? vW : 0.f, hN, hN ? nj.x : 0.f, hN ? vp.x : 0.f, hS, hS ? nm.x : 0.f,
                                   hS ? vm.x : 0.f);
                } else {
                    const float vE = lane_down1(vc.x);
                    z.y = rb_relax(bj.y, cj.y, dj.y, omega, vc.y, hE1, hE1 ? ej.y : 0.f, hE1 ? vE : 0.f, true,
                                   ej.x, vc.x, hN, hN ? nj.y : 0.f, hN ? vp.y : 0.f, hS, hS ? nm.y : 0.f,
                                   hS ? vm.y : 0.f);
                }
                if (own) {
                    typename Pair<TO>::type zv;
                    zv.x = (TO)z.x;
                    zv.y = (TO)z.y;
                    *(typename Pair<TO>::type*)(out + cidx(ix, t, nx)) = zv;
                    acc += (double)bt.x * (double)z.x;
                    acc += (double)bt.y * (double)z.y;
                }
            }
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// Red-black legs with NU sweeps each way on the levels that reach the next
// one by the linear transfers, two rows per wave (k_mg_down_rb2s /
// k_mg_up_rb2s take one).  These levels are latency-bound: a wave's loads are
// issued at once and every sweep is formed in registers, so the extra sweep
// costs arithmetic and wider loads, not another launch.  Each sweep beyond
// the first needs two more rows above and below and one more halo lane on
// either side: a value that a half-sweep forms is wrong one column further in
// from the wave's edge than the values it reads, and the down-leg's
// restriction reads one column beyond a lane's pair, the up-leg's
// prolongation one.  Every cell is formed with the arithmetic of the one-sweep
// legs in the same order; sweeps after the first relax with rb_relax.
template <int NU>
struct RBN {
    static constexpr int G = NU - 1;              // sweeps beyond the first
    static constexpr int HALO = 1 + NU;           // lanes on either side
    static constexpr int OWN = 2 * (WAVE - 2 * HALO);
};
// Relax the red (red_pass) or black cells of register row q (grid row j) of
// U from its neighbours in rows q - 1 and q + 1 and the lanes beside it;
// the pair is (ix, ix + 1) with ix even, so row j's red cell is .x for even j.
template <int NR>
__device__ __forceinline__ void rbn_relax_row(f2v (&U)[NR], int q, int j, bool red_pass, int ny, float omega,
                                              bool hW0, bool hE1, f2v bj, f2v cj, f2v dj, f2v ej, f2v nj, f2v nm) {
    const f2v um = U[q - 1], uc = U[q], up = U[q + 1];
    const bool hN = j < ny - 1, hS = j > 0;
    if (((j & 1) == 0) == red_pass) {
        const float uW = lane_up1(uc.y), aW = lane_up1(ej.y);
        U[q].x = rb_relax(bj.x, cj.x, dj.x, omega, uc.x, true, ej.x, uc.y, hW0, hW0 ? aW : 0.f, hW0 ? uW : 0.f, hN,
                          hN ? nj.x : 0.f, hN ? up.x : 0.f, hS, hS ? nm.x : 0.f, hS ? um.x : 0.f);
    } else {
        const float uE = lane_down1(uc.x);
        U[q].y = rb_relax(bj.y, cj.y, dj.y, omega, uc.y, hE1, hE1 ? ej.y : 0.f, hE1 ? uE : 0.f, true, ej.x, uc.x, hN,
                          hN ? nj.y : 0.f, hN ? up.y : 0.f, hS, hS ? nm.y : 0.f, hS ? um.y : 0.f);
    }
}
// Down-leg for the R rows iy0 .. iy0+R-1 of a wave: from x = 0, NU times red
// then black, then the residual of rows iy0 - 1 .. iy0 + R restricted into
// coarse rows iy0 / 2 .. iy0 / 2 + R / 2 - 1.  Register row q is grid row
// jb + q, jb = iy0 - 3 - 2G: b and D of rows 0 .. NB-1, n of rows 0 .. NB-2,
// e of rows 1 .. NB-2, c of rows 2 .. NB-3.
template <class TB, int NU, int R>
__global__ __launch_bounds__(256) void k_mg_down_rbn(int nx, int ny, int nxc, float omega,
                                                     const float* __restrict__ c, const float* __restrict__ e,
                                                     const float* __restrict__ n, const float* __restrict__ D,
                                                     const TB* __restrict__ b, float* __restrict__ x,
                                                     float* __restrict__ bc, const Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    constexpr int G = RBN<NU>::G, HALO = RBN<NU>::HALO, OWN = RBN<NU>::OWN;
    constexpr int NB = R + 6 + 4 * G;
    static_assert(R == 2 || R == 4, "two or four rows per wave");
    const int done = sc->done;  // arrives with the loads below
    int strip, iy0;
    wide_tile(R, strip, iy0);  // iy0 even
    const int lane = threadIdx.x;
    const int ix = strip * OWN - 2 * HALO + 2 * lane;
    if (iy0 >= ny || strip * OWN >= nx) return;  // uniform over the wave; no barriers follow
    const bool own = lane >= HALO && lane <= WAVE - 1 - HALO && ix < nx;
    const bool inx = ix >= 0 && ix < nx;
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
    const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;
    const f2v z2 = {0.f, 0.f};
    const int jb = iy0 - 3 - 2 * G;
    auto row = [&](int q) -> size_t { return cidx(ixc, min(max(jb + q, 0), ny - 1), nx); };
    f2v Bv[NB], Dv[NB], Nv[NB], Ev[NB], Cv[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const TB2 v = *(const TB2*)(b + row(q));
        Bv[q] = f2v{(float)v.x, (float)v.y};
        Dv[q] = ldc2(D, row(q));
        Nv[q] = q <= NB - 2 ? ldc2(n, row(q)) : z2;
        Ev[q] = (q >= 1 && q <= NB - 2) ? ldc2(e, row(q)) : z2;
        Cv[q] = (q >= 2 && q <= NB - 3) ? ldc2(c, row(q)) : z2;
    }
    // every load is issued before the flag decides
#pragma unroll
    for (int q = 0; q < NB; ++q)
        asm volatile("" : : "v"(Bv[q].x), "v"(Bv[q].y), "v"(Dv[q].x), "v"(Dv[q].y), "v"(Nv[q].x), "v"(Nv[q].y),
                     "v"(Ev[q].x), "v"(Ev[q].y), "v"(Cv[q].x), "v"(Cv[q].y));
    if (done) return;
    // the first sweep's red cells, from zero
    f2v X[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int j = jb + q;
        X[q] = z2;
        if (j >= 0 && j < ny) {
            if ((j & 1) == 0) X[q].x = mg_div(omega * Bv[q].x, Dv[q].x);
            else X[q].y = mg_div(omega * Bv[q].y, Dv[q].y);
        } else {
            Bv[q] = z2;
            Dv[q] = z2;
        }
    }
    // its black cells (rows 1 .. NB-2) from their red neighbours
#pragma unroll
    for (int q = 1; q <= NB - 2; ++q) {
        const int j = jb + q;
        if (j >= 0 && j < ny) {  // uniform over the wave
            const f2v rm = X[q - 1], rc = X[q], rp = X[q + 1], ej = Ev[q], nj = Nv[q], nm = Nv[q - 1];
            const bool hN = j < ny - 1, hS = j > 0;
            if ((j & 1) == 0) {
                const float xE = lane_down1(rc.x);
                X[q].y = rb_black0(Bv[q].y, Dv[q].y, omega, ej.x, rc.x, hE1 ? ej.y : 0.f, hE1 ? xE : 0.f,
                                   hN ? nj.y : 0.f, hN ? rp.y : 0.f, hS ? nm.y : 0.f, hS ? rm.y : 0.f);
            } else {
                const float xW = lane_up1(rc.y), aW = lane_up1(ej.y);
                X[q].x = rb_black0(Bv[q].x, Dv[q].x, omega, hW0 ? aW : 0.f, hW0 ? xW : 0.f, ej.x, rc.y,
                                   hN ? nj.x : 0.f, hN ? rp.x : 0.f, hS ? nm.x : 0.f, hS ? rm.x : 0.f);
            }
        }
    }
    // further sweeps: red cells of rows 2g .. NB-1-2g, then black cells of rows 2g+1 .. NB-2-2g
#pragma unroll
    for (int g = 1; g <= G; ++g) {
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int q = 2 * g + pass; q <= NB - 1 - 2 * g - pass; ++q) {
                const int j = jb + q;
                if (j >= 0 && j < ny)
                    rbn_relax_row(X, q, j, pass == 0, ny, omega, hW0, hE1, Bv[q], Cv[q], Dv[q], Ev[q], Nv[q],
                                  Nv[q - 1]);
            }
        }
    }
    // the residual of rows iy0-1 .. iy0+R (register rows 2+2G .. 3+2G+R),
    // restricted: coarse row iy0/2 + k takes fine rows iy0+2k-1 .. iy0+2k+2
    float acc[R / 2];
#pragma unroll
    for (int k = 0; k < R / 2; ++k) acc[k] = 0.f;
    const float wx0 = ix == 0 ? 1.f : 0.75f, wx1 = ix + 1 == nx - 1 ? 1.f : 0.75f;
#pragma unroll
    for (int s2 = 0; s2 < R + 2; ++s2) {
        const int q = 2 + 2 * G + s2, t = iy0 - 1 + s2;
        if (t >= 0 && t < ny) {  // uniform over the wave
            const f2v X0 = X[q], Xm = X[q - 1], X1 = X[q + 1];
            const f2v e0 = Ev[q], n0 = Nv[q], nm = Nv[q - 1], cc = Cv[q], b0 = Bv[q];
            const float ue = lane_down1(X0.x);
            const float uw = lane_up1(X0.y);
            const float ew = lane_up1(e0.y);
            float a0 = cc.x * X0.x;
            a0 += e0.x * (X0.x - X0.y);
            if (hW0) a0 += ew * (X0.x - uw);
            if (t < ny - 1) a0 += n0.x * (X0.x - X1.x);
            if (t > 0) a0 += nm.x * (X0.x - Xm.x);
            float a1 = cc.y * X0.y;
            if (hE1) a1 += e0.y * (X0.y - ue);
            a1 += e0.x * (X0.y - X0.x);
            if (t < ny - 1) a1 += n0.y * (X0.y - X1.y);
            if (t > 0) a1 += nm.y * (X0.y - Xm.y);
            // P^T: this row's residuals restricted along x (columns ix-1 ..
            // ix+2), weighted into coarse row iy0 / 2
            const float q0 = inx ? b0.x - a0 : 0.f, q1 = inx ? b0.y - a1 : 0.f;
            const float rl = lane_up1(q1), rr = lane_down1(q0);
            float h = 0.25f * rl;
            h += wx0 * q0;
            h += wx1 * q1;
            h += 0.25f * rr;
#pragma unroll
            for (int k = 0; k < R / 2; ++k)
                if (s2 >= 2 * k && s2 <= 2 * k + 3) acc[k] += lin_rw(t, s2 - 2 * k, ny) * h;
            if (own && t >= iy0 && t < iy0 + R) *(f2v*)(x + cidx(ix, t, nx)) = X0;
        }
    }
    const int nyc = (ny + 1) >> 1;
#pragma unroll
    for (int k = 0; k < R / 2; ++k)
        if (own && (iy0 >> 1) + k < nyc) bc[(size_t)(ix >> 1) + (size_t)((iy0 >> 1) + k) * nxc] = acc[k];
}

// Up-leg for the R rows iy0 .. iy0+R-1 of a wave: u = x + scale P xc on rows
// iy0-2-2G .. iy0+R+1+2G (register row q is grid row iy0-2-2G+q), then NU
// times black then red, the last red pass on the wave's rows; b, c, D, e of
// rows 1 .. NR-2, n of rows 0 .. NR-2, xc of coarse rows
// iy0/2-2-G .. iy0/2+R/2+1+G.
template <class TB, class TO, int NU, int R>
__global__ __launch_bounds__(256) void k_mg_up_rbn(int nx, int ny, int nxc, float omega, float scale,
                                                   const float* __restrict__ c, const float* __restrict__ e,
                                                   const float* __restrict__ n, const float* __restrict__ D,
                                                   const TB* __restrict__ b, const float* __restrict__ x,
                                                   const float* __restrict__ xc, TO* __restrict__ out, int with_dot,
                                                   double* partials, Scalars* __restrict__ sc) {
    typedef typename Pair<TB>::type TB2;
    constexpr int G = RBN<NU>::G, HALO = RBN<NU>::HALO, OWN = RBN<NU>::OWN;
    constexpr int NR = R + 4 + 4 * G, NL = R / 2 + 4 + 2 * G;
    static_assert(R == 2 || R == 4, "two or four rows per wave");
    const int done = sc->done;  // arrives with the loads below
    int strip, iy0;
    wide_tile(R, strip, iy0);
    const int lane = threadIdx.x;
    const int ix = strip * OWN - 2 * HALO + 2 * lane;
    double acc = 0.0;
    const bool active = iy0 < ny && strip * OWN < nx;  // uniform over the wave
    const int ixc = ix < 0 ? 0 : (ix < nx ? ix : nx - 2);
    const int iyc = active ? iy0 : 0;  // rows of an inactive wave: any in-grid rows (never used)
    const int jb = iyc - 2 - 2 * G;
    auto row = [&](int q) -> size_t { return cidx(ixc, min(max(jb + q, 0), ny - 1), nx); };
    const int nyc = (ny + 1) >> 1;
    f2v Xv[NR], Cv[NR], Dv[NR], Ev[NR], Nv[NR];
    TB2 Bt[NR];
    float Lc[NL];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        Xv[q] = *(const f2v*)(x + row(q));
        const bool mid = q >= 1 && q <= NR - 2;
        Bt[q] = mid ? *(const TB2*)(b + row(q)) : TB2{};
        Cv[q] = mid ? ldc2(c, row(q)) : f2v{0.f, 0.f};
        Dv[q] = mid ? ldc2(D, row(q)) : f2v{0.f, 0.f};
        Ev[q] = mid ? ldc2(e, row(q)) : f2v{0.f, 0.f};
        Nv[q] = q <= NR - 2 ? ldc2(n, row(q)) : f2v{0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        const int J = min(max((iyc >> 1) - 2 - G + q, 0), nyc - 1);
        Lc[q] = xc[(size_t)(ixc >> 1) + (size_t)J * nxc];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
        asm volatile("" : : "v"(Xv[q].x), "v"(Xv[q].y), "v"(Cv[q].x), "v"(Cv[q].y), "v"(Dv[q].x), "v"(Dv[q].y),
                     "v"(Ev[q].x), "v"(Ev[q].y), "v"(Nv[q].x), "v"(Nv[q].y));
#pragma unroll
    for (int q = 0; q < NL; ++q) asm volatile("" : : "v"(Lc[q]));
    if (done) return;  // uniform over the workgroup, before the partial's barriers
    if (active) {
        const bool own = lane >= HALO && lane <= WAVE - 1 - HALO && ix < nx;
        const bool hW0 = ix > 0, hE1 = ix + 1 < nx - 1;
        const f2v z2 = {0.f, 0.f};
        // u = x + scale P xc: fine row j's own coarse row is Lc[1 + q/2], its
        // neighbour Lc[q/2] (even j) or Lc[2 + q/2] (odd j), folded at the edges
        const float wl = ix == 0 ? 0.f : 0.25f, wr = ix + 1 == nx - 1 ? 0.f : 0.25f;
        f2v U[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int j = jb + q;
            const int i0 = 1 + (q >> 1), i1 = (q & 1) ? i0 + 1 : i0 - 1;
            const int J1 = (iy0 >> 1) - 2 - G + i1;
            const bool fold = J1 < 0 || J1 > nyc - 1;
            const float v = fold ? Lc[i0] : 0.75f * Lc[i0] + 0.25f * Lc[i1];
            const float vl = lane_up1(v), vr = lane_down1(v);
            const float p0 = (1.f - wl) * v + wl * vl, p1 = (1.f - wr) * v + wr * vr;
            U[q] = (j >= 0 && j < ny) ? f2v{Xv[q].x + scale * p0, Xv[q].y + scale * p1} : z2;
        }
        // NU times: black cells of rows 1+2g .. NR-2-2g, red cells of rows 2+2g .. NR-3-2g
#pragma unroll
        for (int g = 0; g <= G; ++g) {
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
                for (int q = 1 + 2 * g + pass; q <= NR - 2 - 2 * g - pass; ++q) {
                    const int j = jb + q;
                    if (j >= 0 && j < ny)
                        rbn_relax_row(U, q, j, pass == 1, ny, omega, hW0, hE1,
                                      f2v{(float)Bt[q].x, (float)Bt[q].y}, Cv[q], Dv[q], Ev[q], Nv[q], Nv[q - 1]);
                }
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < R; ++s2) {
            const int q = 2 + 2 * G + s2, t = iy0 + s2;
            if (own && t < ny) {
                const f2v z = U[q];
                typename Pair<TO>::type zv;
                zv.x = (TO)z.x;
                zv.y = (TO)z.y;
                *(typename Pair<TO>::type*)(out + cidx(ix, t, nx)) = zv;
                acc += (double)Bt[q].x * (double)z.x;
                acc += (double)Bt[q].y * (double)z.y;
            }
        }
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// Red-black legs for the small levels and for levels of odd width: one
// thread per coarse cell (down) or fine cell (up), every neighbour's value
// formed from global memory with the helpers above.
template <class TB>
__global__ __launch_bounds__(256) void k_mg_down_rb_flat(int nx, int ny, int nxc, int nyc, float omega,
                                                         const float* __restrict__ c, const float* __restrict__ e,
                                                         const float* __restrict__ n, const float* __restrict__ D,
                                                         const TB* __restrict__ b, float* __restrict__ x,
                                                         float* __restrict__ bc, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= nxc || J >= nyc) return;
    auto xs = [&](int jx, int jy) -> float { return rb_down_x<TB>(nx, ny, jx, jy, omega, e, n, D, b); };
    float acc = 0.f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int ix = 2 * I + dx, iy = 2 * J + dy;
            if (ix < nx && iy < ny) {
                const size_t k = (size_t)ix + (size_t)iy * nx;
                const float u = xs(ix, iy);
                x[k] = u;
                float a = c[k] * u;
                if (ix < nx - 1) a += e[k] * (u - xs(ix + 1, iy));
                if (ix > 0) a += e[k - 1] * (u - xs(ix - 1, iy));
                if (iy < ny - 1) a += n[k] * (u - xs(ix, iy + 1));
                if (iy > 0) a += n[k - nx] * (u - xs(ix, iy - 1));
                acc += (float)b[k] - a;
            }
        }
    bc[(size_t)I + (size_t)J * nxc] = acc;
}
template <class TB>
__global__ __launch_bounds__(256) void k_mg_up_rb_flat(int nx, int ny, int nxc, float omega, float scale,
                                                       const float* __restrict__ c, const float* __restrict__ e,
                                                       const float* __restrict__ n, const float* __restrict__ D,
                                                       const TB* __restrict__ b, const float* __restrict__ x,
                                                       const float* __restrict__ xc, TB* __restrict__ out,
                                                       int with_dot, double* partials, Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    double acc = 0.0;
    if (ix < nx && iy < ny) {
        auto uat = [&](int jx, int jy) -> float {
            return x[(size_t)jx + (size_t)jy * nx] + scale * xc[(size_t)(jx >> 1) + (size_t)(jy >> 1) * nxc];
        };
        auto vat = [&](int jx, int jy) -> float {  // a black neighbour, relaxed
            return rb_up_black<TB>(nx, ny, jx, jy, omega, c, e, n, D, b, uat);
        };
        const size_t k = (size_t)ix + (size_t)iy * nx;
        float z;
        if ((ix + iy) & 1) {
            z = rb_up_black<TB>(nx, ny, ix, iy, omega, c, e, n, D, b, uat);
        } else {
            const bool hE = ix < nx - 1, hW = ix > 0, hN = iy < ny - 1, hS = iy > 0;
            z = rb_relax((float)b[k], c[k], D[k], omega, uat(ix, iy), hE, hE ? e[k] : 0.f, hE ? vat(ix + 1, iy) : 0.f,
                         hW, hW ? e[k - 1] : 0.f, hW ? vat(ix - 1, iy) : 0.f, hN, hN ? n[k] : 0.f,
                         hN ? vat(ix, iy + 1) : 0.f, hS, hS ? n[k - nx] : 0.f, hS ? vat(ix, iy - 1) : 0.f);
        }
        const TB bk = b[k];
        out[k] = (TB)z;
        acc = (double)bk * (double)z;
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// The same two legs for small levels (at most MG_FLAT_CELLS cells) outside
// the tail kernel, one thread per coarse cell (down) or fine cell (up),
// reading the neighbours' values directly.
template <class TB>
__global__ __launch_bounds__(256) void k_mg_down_flat(int nx, int ny, int nxc, int nyc, float omega,
                                                      const float* __restrict__ c, const float* __restrict__ e,
                                                      const float* __restrict__ n, const float* __restrict__ D,
                                                      const TB* __restrict__ b, float* __restrict__ x,
                                                      float* __restrict__ bc, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= nxc || J >= nyc) return;
    auto xs = [&](size_t k) -> float { return mg_div(omega * (float)b[k], D[k]); };
    float acc = 0.f;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
        float xv[2] = {0.f, 0.f};
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const int ix = 2 * I + dx, iy = 2 * J + dy;
            if (ix < nx && iy < ny) {
                const size_t k = (size_t)ix + (size_t)iy * nx;
                xv[dx] = xs(k);
                acc += (float)b[k] - flux_apply(nx, ny, ix, iy, k, c, e, n, xs);
            }
        }
        // the pair of x values as one 8 B store when it is aligned (even nx),
        // so each cache line is written whole
        const int iy = 2 * J + dy;
        if (iy < ny) {
            const size_t k0 = (size_t)(2 * I) + (size_t)iy * nx;
            if ((nx & 1) == 0) {
                *(float2*)(x + k0) = make_float2(xv[0], xv[1]);
            } else {
                x[k0] = xv[0];
                if (2 * I + 1 < nx) x[k0 + 1] = xv[1];
            }
        }
    }
    bc[(size_t)I + (size_t)J * nxc] = acc;
}
// with_dot: also the partial of b'z (the finest level of FP64 CG's V-cycle)
template <class TB>
__global__ __launch_bounds__(256) void k_mg_up_flat(int nx, int ny, int nxc, float omega, float scale,
                                                    const float* __restrict__ c, const float* __restrict__ e,
                                                    const float* __restrict__ n, const float* __restrict__ D,
                                                    const TB* __restrict__ b, const float* __restrict__ x,
                                                    const float* __restrict__ xc, TB* __restrict__ out,
                                                    int with_dot, double* partials, Scalars* __restrict__ sc) {
    if (sc->done) return;
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y * blockDim.y + threadIdx.y;
    double acc = 0.0;
    if (ix < nx && iy < ny) {
        auto uat = [&](int jx, int jy) -> float {
            return x[(size_t)jx + (size_t)jy * nx] + scale * xc[(size_t)(jx >> 1) + (size_t)(jy >> 1) * nxc];
        };
        const size_t k = (size_t)ix + (size_t)iy * nx;
        const float u = uat(ix, iy);
        float a = c[k] * u;
        if (ix < nx - 1) a += e[k] * (u - uat(ix + 1, iy));
        if (ix > 0) a += e[k - 1] * (u - uat(ix - 1, iy));
        if (iy < ny - 1) a += n[k] * (u - uat(ix, iy + 1));
        if (iy > 0) a += n[k - nx] * (u - uat(ix, iy - 1));
        const TB bk = b[k];
        const float z = u + mg_div(omega * ((float)bk - a), D[k]);
        out[k] = (TB)z;
        acc = (double)bk * (double)z;
    }
    if (with_dot) {
        double v[1] = {acc};
        block_partial<1, false>(v, partials);
        if (with_dot == 2 && lead_thread()) sc->gpend = 1;  // the next CG matvec finishes r'z
    }
}

// k / w for 0 <= k < 2^21 through the FP32 reciprocal of w: (k + 1/2) / w
// lies at least 1/(2w) from every integer and the two roundings move it by a
// relative 2^-23 at most, so truncation gives the quotient exactly (checked
// on the CPU for every w up to 4096, k below 2^16, and for k below 2^21 over a
// spread of widths).
struct FastDiv {
    int w;
    float rw;
    __device__ explicit FastDiv(int w_) : w(w_), rw(1.0f / (float)w_) {}
    __device__ int quot(int k) const { return (int)(((float)k + 0.5f) * rw); }
};

// The small levels of the V-cycle, from the first level of at most
// MG_TAIL_CELLS cells down to the coarsest, in one workgroup: down-legs, the
// dense coarsest solve, up-legs.  Everything lives in LDS (dynamic shared
// memory): every level's b and x (x is overwritten by the level's result
// after a barrier), its c, e, n, D and the coarsest inverse, 6 floats per
// cell plus the inverse.  The first tail level's b comes from global memory
// and its result goes back there.  With at most MG_TAIL_CELLS cells per level
// each thread owns at most 4 cells of a level.
__global__ __launch_bounds__(1024) void k_mg_tail(MGTail T, const Scalars* __restrict__ sc) {
    const int done = sc->done;  // read together with the loads below
    extern __shared__ float4 lds4[];
    float* lds = (float*)lds4;
    __shared__ float red[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const float omega = T.omega, scale = T.scale;
    int off[MG_MAX_TAIL + 1];
    off[0] = 0;
    for (int l = 0; l < T.nlev; ++l) off[l + 1] = off[l] + T.L[l].nx * T.L[l].ny;
    const int tp = T.stride;  // >= off[nlev], a multiple of 4
    const int lc = T.nlev - 1;
    const int ncell = T.L[lc].nx * T.L[lc].ny;
    // LDS: rhs, x, the four coefficient arrays of every tail level, and the
    // coarsest inverse.  The last five come packed from set-up in this very
    // layout and are copied with 16 B loads, all of a thread's loads in
    // flight at once (one memory round trip); every later phase reads LDS.
    float* Bs = lds;           // right-hand sides
    float* Xs = lds + tp;      // smoothed x, then u, then the level's result
    float* Cs = lds + 2 * tp;  // reaction c
    float* Es = lds + 3 * tp;  // east couplings
    float* Ns = lds + 4 * tp;  // north couplings
    float* Ds = lds + 5 * tp;  // Jacobi diagonal
    float* Is = lds + 6 * tp;  // coarsest inverse, ncell x ncell
    float* Hs = lds + 2 * tp + ((4 * tp + ncell * ncell + 3) & ~3);  // linear restriction: x-restricted rows
    {
        constexpr int UB = 4, UP = 8;
        const float* b0 = T.L[0].b;
        const int nb = off[1];
        float vb[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) vb[u] = b0[min(tid + u * nt, nb - 1)];
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4* src = (const f4*)T.pack;
        f4* dst = (f4*)(lds4 + (2 * tp) / 4);
        const int n4 = (4 * tp + ncell * ncell + 3) / 4;
        // unconditional loads (index clamped into the block), all issued
        // before any of them is used: the empty asm statements take every
        // loaded value, so the compiler cannot move a load down to its
        // (conditional) store; the done flag arrives with them
        for (int base = tid;; base += UP * nt) {
            f4 v[UP];
#pragma unroll
            for (int u = 0; u < UP; ++u) v[u] = src[min(base + u * nt, n4 - 1)];
#pragma unroll
            for (int u = 0; u < UB; ++u) asm volatile("" : : "v"(vb[u]));
#pragma unroll
            for (int u = 0; u < UP; ++u) asm volatile("" : : "v"(v[u]));
            if (done) return;  // uniform: every thread leaves here, before any barrier
#pragma unroll
            for (int u = 0; u < UP; ++u) {
                const int i = base + u * nt;
                if (i < n4) dst[i] = v[u];
            }
            if (base + UP * nt >= n4) break;
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int k = tid + u * nt;
            if (k < nb) Bs[k] = vb[u];
        }
        for (int k = tid + UB * nt; k < nb; k += nt) Bs[k] = b0[k];
    }
    __syncthreads();
    for (int l = 0; l + 1 < T.nlev; ++l) {
        const int fnx = T.L[l].nx, fny = T.L[l].ny, cnx = T.L[l + 1].nx, cny = T.L[l + 1].ny;
        const float* b = Bs + off[l];
        const float *c = Cs + off[l], *e = Es + off[l], *n = Ns + off[l], *D = Ds + off[l];
        float* x = Xs + off[l];
        float* bc = Bs + off[l + 1];
        auto xs = [&](size_t k) -> float { return mg_div(omega * b[k], D[k]); };
        const FastDiv dfx(fnx), dcx(cnx);
        if (T.rb) {
            if ((fnx & 1) == 0) {
                // even width: each pass takes only its own colour's cells, hx per
                // row (red ix = 2 j + (iy & 1), black the other); red x, then black
                // from the stored red values
                const int hx = fnx >> 1;
                const FastDiv dh(hx);
                for (int p = tid; p < hx * fny; p += nt) {
                    const int iy = dh.quot(p), k = 2 * (p - iy * hx) + (iy & 1) + iy * fnx;
                    x[k] = xs(k);
                }
                __syncthreads();
                for (int p = tid; p < hx * fny; p += nt) {
                    const int iy = dh.quot(p), ix = 2 * (p - iy * hx) + ((iy + 1) & 1), k = ix + iy * fnx;
                    const bool hW = ix > 0, hE = ix < fnx - 1, hN = iy < fny - 1, hS = iy > 0;
                    x[k] = rb_black0(b[k], D[k], omega, hW ? e[k - 1] : 0.f, hW ? x[k - 1] : 0.f, hE ? e[k] : 0.f,
                                     hE ? x[k + 1] : 0.f, hN ? n[k] : 0.f, hN ? x[k + fnx] : 0.f,
                                     hS ? n[k - fnx] : 0.f, hS ? x[k - fnx] : 0.f);
                }
            } else {
                // red-black: every cell's x (black from its red neighbours' values, recomputed)
                for (int k = tid; k < fnx * fny; k += nt) {
                    const int iy = dfx.quot(k), ix = k - iy * fnx;
                    if (((ix + iy) & 1) == 0) {
                        x[k] = xs(k);
                    } else {
                        const bool hW = ix > 0, hE = ix < fnx - 1, hN = iy < fny - 1, hS = iy > 0;
                        x[k] = rb_black0(b[k], D[k], omega, hW ? e[k - 1] : 0.f, hW ? xs(k - 1) : 0.f,
                                         hE ? e[k] : 0.f, hE ? xs(k + 1) : 0.f, hN ? n[k] : 0.f,
                                         hN ? xs(k + fnx) : 0.f, hS ? n[k - fnx] : 0.f, hS ? xs(k - fnx) : 0.f);
                    }
                }
            }
            __syncthreads();
            auto xk = [&](size_t k) -> float { return x[k]; };
            if ((T.lin_mask >> l) & 1) {
                // P^T (b - A x): each fine row restricted along x into Hs
                // (cnx x fny), then the rows combined with the y weights.
                // Thread q = (I, iy) forms the residuals of fine cells 2I and
                // 2I+1; those of 2I-1 and 2I+2 come from the lanes of I-1 and
                // I+1 (q-1, q+1), or are formed here at a wave's edge.  Each
                // term is the same value, added in the same order, as when
                // every thread formed its four.  Every lane runs every trip,
                // so the lane shifts read active lanes.
                auto res = [&](int ix, int iy) -> float {
                    const size_t k = (size_t)ix + (size_t)iy * fnx;
                    return b[k] - flux_apply(fnx, fny, ix, iy, k, c, e, n, xk);
                };
                const int nq = cnx * fny;
                for (int q0 = 0; q0 < nq; q0 += nt) {
                    const int q = q0 + tid;
                    const bool act = q < nq;
                    const int qc = act ? q : nq - 1;
                    const int iy = dcx.quot(qc), I = qc - iy * cnx;
                    const int lane = tid & (WAVE - 1);
                    const float ra = res(2 * I, iy);
                    const float rb = 2 * I + 1 < fnx ? res(2 * I + 1, iy) : 0.f;
                    float rm = lane_up1(rb), rp = lane_down1(ra);
                    const float w0 = lin_rw(2 * I - 1, 0, fnx), w1 = lin_rw(2 * I, 1, fnx);
                    const float w2 = lin_rw(2 * I + 1, 2, fnx), w3 = lin_rw(2 * I + 2, 3, fnx);
                    if (w0 != 0.f && (lane == 0 || q == q0)) rm = res(2 * I - 1, iy);
                    if (w3 != 0.f && (lane == WAVE - 1 || qc != q)) rp = res(2 * I + 2, iy);
                    float h = 0.f;
                    if (w0 != 0.f) h += w0 * rm;
                    if (w1 != 0.f) h += w1 * ra;
                    if (w2 != 0.f) h += w2 * rb;
                    if (w3 != 0.f) h += w3 * rp;
                    if (act) Hs[q] = h;
                }
                __syncthreads();
                for (int q = tid; q < cnx * cny; q += nt) {
                    const int J = dcx.quot(q), I = q - J * cnx;
                    float acc = 0.f;
#pragma unroll
                    for (int p = 0; p < 4; ++p) {
                        const int iy = 2 * J - 1 + p;
                        const float wy = lin_rw(iy, p, fny);
                        if (wy != 0.f) acc += wy * Hs[I + iy * cnx];
                    }
                    bc[q] = acc;
                }
                __syncthreads();
                continue;
            }
            for (int q = tid; q < cnx * cny; q += nt) {
                const int J = dcx.quot(q), I = q - J * cnx;
                float acc = 0.f;
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const int ix = 2 * I + dx, iy = 2 * J + dy;
                        if (ix < fnx && iy < fny) {
                            const size_t k = (size_t)ix + (size_t)iy * fnx;
                            acc += b[k] - flux_apply(fnx, fny, ix, iy, k, c, e, n, xk);
                        }
                    }
                bc[q] = acc;
            }
            __syncthreads();
            continue;
        }
        for (int q = tid; q < cnx * cny; q += nt) {
            const int J = dcx.quot(q), I = q - J * cnx;
            float acc = 0.f;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const int ix = 2 * I + dx, iy = 2 * J + dy;
                    if (ix < fnx && iy < fny) {
                        const size_t k = (size_t)ix + (size_t)iy * fnx;
                        x[k] = xs(k);
                        acc += b[k] - flux_apply(fnx, fny, ix, iy, k, c, e, n, xs);
                    }
                }
            bc[q] = acc;
        }
        __syncthreads();
    }
    {
        // coarsest: x = inv b with the symmetric inverse read by columns;
        // 16 threads per row (one per wave, the lanes of a wave on 64
        // consecutive rows, so a wave reads 64 consecutive floats), 4
        // products each, the 16 parts summed in a fixed order
        const float* b = Bs + off[lc];
        const int i = tid & (WAVE - 1), part = tid >> 6;
        float acc = 0.f;
        if (i < ncell)
            for (int j = part * 4; j < part * 4 + 4 && j < ncell; ++j) acc += Is[j * ncell + i] * b[j];
        red[part * WAVE + i] = acc;
        __syncthreads();
        if (part == 0 && i < ncell) {
            float s = 0.f;
            for (int q = 0; q < 16; ++q) s += red[q * WAVE + i];
            Xs[off[lc] + i] = s;
        }
        __syncthreads();
    }
    for (int l = T.nlev - 2; l >= 0; --l) {
        const int fnx = T.L[l].nx, fny = T.L[l].ny, cnx = T.L[l + 1].nx;
        const int nf = fnx * fny;
        const float* b = Bs + off[l];
        const float *c = Cs + off[l], *e = Es + off[l], *n = Ns + off[l], *D = Ds + off[l];
        float* x = Xs + off[l];
        const float* zc = Xs + off[l + 1];
        const FastDiv dfx(fnx);
        const bool lin = (T.lin_mask >> l) & 1;
        if (T.rb && (fnx & 1) == 0 && !lin) {
            // even width: u = x + scale P xc formed where it is read, black cells
            // relaxing from u, then red cells (own u, relaxed black neighbours);
            // each pass takes only its own colour's cells
            const int hx = fnx >> 1;
            const FastDiv dh(hx);
            auto uat = [&](int k, int ix, int iy) -> float { return x[k] + scale * zc[(ix >> 1) + (iy >> 1) * cnx]; };
            for (int p = tid; p < hx * fny; p += nt) {
                const int iy = dh.quot(p), ix = 2 * (p - iy * hx) + ((iy + 1) & 1), k = ix + iy * fnx;
                const bool hE = ix < fnx - 1, hW = ix > 0, hN = iy < fny - 1, hS = iy > 0;
                x[k] = rb_relax(b[k], c[k], D[k], omega, uat(k, ix, iy), hE, hE ? e[k] : 0.f,
                                hE ? uat(k + 1, ix + 1, iy) : 0.f, hW, hW ? e[k - 1] : 0.f,
                                hW ? uat(k - 1, ix - 1, iy) : 0.f, hN, hN ? n[k] : 0.f,
                                hN ? uat(k + fnx, ix, iy + 1) : 0.f, hS, hS ? n[k - fnx] : 0.f,
                                hS ? uat(k - fnx, ix, iy - 1) : 0.f);
            }
            __syncthreads();
            for (int p = tid; p < hx * fny; p += nt) {
                const int iy = dh.quot(p), ix = 2 * (p - iy * hx) + (iy & 1), k = ix + iy * fnx;
                const bool hE = ix < fnx - 1, hW = ix > 0, hN = iy < fny - 1, hS = iy > 0;
                x[k] = rb_relax(b[k], c[k], D[k], omega, uat(k, ix, iy), hE, hE ? e[k] : 0.f, hE ? x[k + 1] : 0.f,
                                hW, hW ? e[k - 1] : 0.f, hW ? x[k - 1] : 0.f, hN, hN ? n[k] : 0.f,
                                hN ? x[k + fnx] : 0.f, hS, hS ? n[k - fnx] : 0.f, hS ? x[k - fnx] : 0.f);
            }
            __syncthreads();
            continue;
        }
        if (lin) {
            for (int k = tid; k < nf; k += nt) {
                const int iy = dfx.quot(k), ix = k - iy * fnx;
                const LinP px = lin_p(ix, fnx), py = lin_p(iy, fny);
                const float v0 = px.w0 * zc[px.I0 + py.I0 * cnx] + px.w1 * zc[px.I1 + py.I0 * cnx];
                const float v1 = px.w0 * zc[px.I0 + py.I1 * cnx] + px.w1 * zc[px.I1 + py.I1 * cnx];
                x[k] = x[k] + T.scale_lin * (py.w0 * v0 + py.w1 * v1);
            }
        } else {
            for (int k = tid; k < nf; k += nt) {
                const int iy = dfx.quot(k), ix = k - iy * fnx;
                x[k] = x[k] + scale * zc[(ix >> 1) + (iy >> 1) * cnx];
            }
        }
        __syncthreads();
        auto us = [&](size_t k) -> float { return x[k]; };
        if (T.rb) {
            // black cells relax from u in place (they read only red neighbours),
            // then red cells from the relaxed black ones; on even widths each
            // pass takes only its own colour's cells (hx per row)
            auto relax_at = [&](int k, int ix, int iy) {
                const bool hE = ix < fnx - 1, hW = ix > 0, hN = iy < fny - 1, hS = iy > 0;
                x[k] = rb_relax(b[k], c[k], D[k], omega, x[k], hE, hE ? e[k] : 0.f, hE ? x[k + 1] : 0.f, hW,
                                hW ? e[k - 1] : 0.f, hW ? x[k - 1] : 0.f, hN, hN ? n[k] : 0.f,
                                hN ? x[k + fnx] : 0.f, hS, hS ? n[k - fnx] : 0.f, hS ? x[k - fnx] : 0.f);
            };
            const int hx = fnx >> 1;
            const FastDiv dh(hx > 0 ? hx : 1);
            for (int colour = 1; colour >= 0; --colour) {
                if ((fnx & 1) == 0) {
                    for (int p = tid; p < hx * fny; p += nt) {
                        const int iy = dh.quot(p), ix = 2 * (p - iy * hx) + ((iy + colour) & 1);
                        relax_at(ix + iy * fnx, ix, iy);
                    }
                } else {
                    for (int k = tid; k < nf; k += nt) {
                        const int iy = dfx.quot(k), ix = k - iy * fnx;
                        if (((ix + iy) & 1) == colour) relax_at(k, ix, iy);
                    }
                }
                __syncthreads();
            }
            continue;
        }
        float z[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = tid + j * nt;
            if (k < nf) {
                const int iy = dfx.quot(k), ix = k - iy * fnx;
                z[j] = x[k] + mg_div(omega * (b[k] - flux_apply(fnx, fny, ix, iy, (size_t)k, c, e, n, us)), D[k]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = tid + j * nt;
            if (k < nf) x[k] = z[j];
        }
        __syncthreads();
    }
    for (int k = tid; k < off[1]; k += nt) T.L[0].xb[k] = Xs[k];
}

// Coarsest level (at most 64 cells): assemble the dense FP64 matrix and invert
// it in place (Gauss-Jordan without pivoting, the matrix being SPD); the
// inverse is stored in FP32:
//   p = A_kk;  A_kk = 1/p;  A_kj /= p (j != k);  A_ij -= A_ik A_kj (i, j != k);
//   A_ik = -A_ik / p (i != k).
// Workgroup 64 x 16: thread (j, ty) keeps A_ij for the rows i = ty + 16 m
// (m = 0..3) in registers, so each wave holds all 64 columns of its rows and
// takes the pivot column A_ik from lane k by readlane; the pivot row goes
// through LDS, double-buffered, with one barrier a step.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long w = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)w, lane);
    const int hi = __builtin_amdgcn_readlane((int)(w >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__global__ __launch_bounds__(1024) void k_mg_dense_inverse_lds(int nx, int ny, const float* __restrict__ c,
                                                               const float* __restrict__ e,
                                                               const float* __restrict__ nn,
                                                               float* __restrict__ inv) {
    constexpr int NM = 64;  // the largest coarsest level, and the LDS pitch
    __shared__ double A[NM * NM];
    __shared__ double rowk[2][NM];
    const int n = nx * ny;
    const int j = threadIdx.x, ty = threadIdx.y;
    const int tid = j + WAVE * ty;
    for (int q = tid; q < NM * NM; q += WAVE * 16) A[q] = 0.0;
    __syncthreads();
    if (tid < n) {
        const int k = tid;
        const int ix = k % nx, iy = k / nx;
        double d = c[k];
        if (ix < nx - 1) { d += e[k]; A[k * NM + k + 1] = -(double)e[k]; }
        if (ix > 0) { d += e[k - 1]; A[k * NM + k - 1] = -(double)e[k - 1]; }
        if (iy < ny - 1) { d += nn[k]; A[k * NM + k + nx] = -(double)nn[k]; }
        if (iy > 0) { d += nn[k - nx]; A[k * NM + k - nx] = -(double)nn[k - nx]; }
        A[k * NM + k] = d;
    }
    __syncthreads();
    double a[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int i = ty + 16 * m;
        a[m] = (i < n && j < n) ? A[i * NM + j] : 0.0;
    }
    for (int k = 0; k < n; ++k) {
        double* rb = rowk[k & 1];
        if (ty == (k & 15)) {  // the pivot row's wave (uniform branch)
            const int q = k >> 4;
            rb[j] = q == 0 ? a[0] : (q == 1 ? a[1] : (q == 2 ? a[2] : a[3]));
        }
        __syncthreads();  // the other buffer is rewritten only after the next barrier
        const double ip = 1.0 / rb[k];
        const double akj = (j == k) ? ip : rb[j] * ip;  // the scaled pivot row
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int i = ty + 16 * m;
            const double cik = readlane_f64(a[m], k);  // A_ik before this step
            if (i < n && j < n) a[m] = (i == k) ? akj : ((j == k) ? -cik * ip : a[m] - cik * akj);
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int i = ty + 16 * m;
        if (i < n && j < n) inv[i * n + j] = (float)a[m];
    }
}


// ------------------------------------------------ learned warm start (FP64)
// The network sees the Picard iteration's system Jacobi-scaled: with
// D = diag(aP), r^ = D^-1/2 r / s and A^ = D^-1/2 A D^-1/2 (unit diagonal,
// couplings aF / sqrt(aP aP_nb)), the correction e solves A^ (D^1/2 e / s) = r^.
// The image is the grid padded with zero cells to W8 x H8.

typedef __bf16 nn_bf16x8 __attribute__((ext_vector_type(8)));

// r = rhs - Ap and the partials of sum r^2 / aP; grid-stride over n cells.
__global__ __launch_bounds__(256) void k_nn_resid(size_t n, const double* __restrict__ aP,
                                                  const double* __restrict__ rhs, const double* __restrict__ Ap,
                                                  double* __restrict__ r, double* partials) {
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * WAVE * WAVES + tid; i < n; i += stride) {
        const double v = rhs[i] - Ap[i];
        r[i] = v;
        acc += v * v / aP[i];
    }
    double a[1] = {acc};
    block_partial<1, false>(a, partials);
}

// Network input (8 BF16 channels per pixel): r^, 2x the four Jacobi-scaled
// couplings (E, W, N, S), c/aP, 0, 0; zero outside the grid.
__global__ __launch_bounds__(256) void k_nn_features(int nx, int ny, int W8, int H8, const double* __restrict__ aP,
                                                     const double* __restrict__ aE, const double* __restrict__ aN,
                                                     const double* __restrict__ cP, const double* __restrict__ r,
                                                     const Scalars* __restrict__ sc, int slot,
                                                     nn_bf16x8* __restrict__ x) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)W8 * H8) return;
    const int ix = (int)(p % W8), iy = (int)(p / W8);
    nn_bf16x8 o = {};
    if (ix < nx && iy < ny) {
        const size_t k = (size_t)ix + (size_t)iy * nx;
        const double d = aP[k], sd = sqrt(d);
        const double s = slot ? sc->nn_scale_pc : sc->nn_scale;
        o[0] = (__bf16)(s > 0.0 ? r[k] / (sd * s) : 0.0);
        o[1] = (__bf16)(ix < nx - 1 ? 2.0 * aE[k] / sqrt(d * aP[k + 1]) : 0.0);
        o[2] = (__bf16)(ix > 0 ? 2.0 * aE[k - 1] / sqrt(d * aP[k - 1]) : 0.0);
        o[3] = (__bf16)(iy < ny - 1 ? 2.0 * aN[k] / sqrt(d * aP[k + nx]) : 0.0);
        o[4] = (__bf16)(iy > 0 ? 2.0 * aN[k - nx] / sqrt(d * aP[k - nx]) : 0.0);
        o[5] = (__bf16)(cP[k] / d);
    }
    x[p] = o;
}

// e = s N / sqrt(aP) on cells with at least one coupling (0 on identity and
// isolated rows, whose solution the warm start must not move).  slot 0 takes
// the warm start's scale, 1 the preconditioner's.
__global__ __launch_bounds__(256) void k_nn_correction(int nx, int ny, int W8, const float* __restrict__ N,
                                                       const double* __restrict__ aP, const double* __restrict__ aE,
                                                       const double* __restrict__ aN, const Scalars* __restrict__ sc,
                                                       int slot, double* __restrict__ e) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (size_t)nx * ny) return;
    const int ix = (int)(k % nx), iy = (int)(k / nx);
    const bool coupled =
        aE[k] != 0.0 || aN[k] != 0.0 || (ix > 0 && aE[k - 1] != 0.0) || (iy > 0 && aN[k - nx] != 0.0);
    const double s = slot ? sc->nn_scale_pc : sc->nn_scale;
    e[k] = coupled ? s * (double)N[(size_t)ix + (size_t)iy * W8] / sqrt(aP[k]) : 0.0;
}

// Partials of e'r and e'Ae for the energy-minimising step.
__global__ __launch_bounds__(256) void k_nn_ls(size_t n, const double* __restrict__ e, const double* __restrict__ r,
                                               const double* __restrict__ Ae, double* partials) {
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    double a0 = 0.0, a1 = 0.0;
    for (size_t i = (size_t)blockIdx.x * WAVE * WAVES + tid; i < n; i += stride) {
        a0 += e[i] * r[i];
        a1 += e[i] * Ae[i];
    }
    double a[2] = {a0, a1};
    block_partial<2, false>(a, partials);
}

// x0 = h + gamma e (x0 = h exactly when the step was rejected).
__global__ __launch_bounds__(256) void k_nn_warm(size_t n, const double* __restrict__ h,
                                                 const double* __restrict__ e, const Scalars* __restrict__ sc,
                                                 double* __restrict__ x0) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double g = sc->nn_gamma;
    x0[i] = g != 0.0 ? h[i] + g * e[i] : h[i];
}

// Training target sqrt(aP) (h_new - h_prev) / s on the padded image.
__global__ __launch_bounds__(256) void k_nn_target(int nx, int ny, int W8, int H8, const double* __restrict__ aP,
                                                   const double* __restrict__ h_new,
                                                   const double* __restrict__ h_prev,
                                                   const Scalars* __restrict__ sc, float* __restrict__ t) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)W8 * H8) return;
    const int ix = (int)(p % W8), iy = (int)(p / W8);
    float v = 0.f;
    const double s = sc->nn_scale;
    if (ix < nx && iy < ny && s > 0.0) {
        const size_t k = (size_t)ix + (size_t)iy * nx;
        v = (float)(sqrt(aP[k]) * (h_new[k] - h_prev[k]) / s);
    }
    t[p] = v;
}

// Training target of a learned CG iteration: sqrt(aP) (x_final - x_k) / s_k,
// the remaining error at iteration k in the network's scaling.
__global__ __launch_bounds__(256) void k_nn_target_cg(int nx, int ny, int W8, int H8, const double* __restrict__ aP,
                                                      const double* __restrict__ x_final,
                                                      const double* __restrict__ x_k,
                                                      const double* __restrict__ s_k, float* __restrict__ t) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (size_t)W8 * H8) return;
    const int ix = (int)(p % W8), iy = (int)(p / W8);
    float v = 0.f;
    const double s = *s_k;
    if (ix < nx && iy < ny && s > 0.0) {
        const size_t k = (size_t)ix + (size_t)iy * nx;
        v = (float)(sqrt(aP[k]) * (x_final[k] - x_k[k]) / s);
    }
    t[p] = v;
}

// Reference norm of the stopping test: sqrt(r'z) of the preconditioner just
// applied to r(h) (mode 0), or the Jacobi norm s sqrt(n) (mode 1).
__global__ void k_nn_ref(Scalars* sc, int mode, double n) {
    if (threadIdx.x != 0) return;
    sc->rnorm0_ref = mode == 0 ? sqrt(sc->gamma) : sc->nn_scale * sqrt(n);
}

// Learned preconditioner inside flexible CG (one iteration in m):
//   s = rms of r / sqrt(aP);  e = network correction for r;  gamma = e'r / e'Ae;
//   z = gamma e + P(r - gamma A e)   (P: the V-cycle or the Chebyshev polynomial).
// A rejected step (gamma = 0) leaves z = P(r).  These kernels return at once
// after convergence, like the CG kernels.
__global__ __launch_bounds__(256) void k_nn_rscale(size_t n, const double* __restrict__ aP,
                                                   const double* __restrict__ r, double* partials,
                                                   const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * WAVE * WAVES + tid; i < n; i += stride) acc += r[i] * r[i] / aP[i];
    double a[1] = {acc};
    block_partial<1, false>(a, partials);
}
__global__ __launch_bounds__(256) void k_nn_r1(size_t n, const double* __restrict__ r, const double* __restrict__ Ae,
                                               const Scalars* __restrict__ sc, double* __restrict__ r1) {
    if (sc->done) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double g = sc->nn_gamma_pc;
    r1[i] = g != 0.0 ? r[i] - g * Ae[i] : r[i];
}
__global__ __launch_bounds__(256) void k_nn_combine(size_t n, const double* __restrict__ r,
                                                    const double* __restrict__ e, double* __restrict__ z,
                                                    double* partials, const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    const double g = sc->nn_gamma_pc;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * WAVE * WAVES + tid; i < n; i += stride) {
        const double zi = g != 0.0 ? z[i] + g * e[i] : z[i];
        z[i] = zi;
        acc += r[i] * zi;
    }
    double a[1] = {acc};
    block_partial<1, false>(a, partials);
}
// Partial of a'b (the flexible-CG beta numerator z'Ap).
__global__ __launch_bounds__(256) void k_dot_partial(size_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* partials,
                                                     const Scalars* __restrict__ sc) {
    if (sc->done) return;
    const size_t tid = (size_t)threadIdx.x + threadIdx.y * WAVE;
    const size_t stride = (size_t)gridDim.x * WAVE * WAVES;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * WAVE * WAVES + tid; i < n; i += stride) acc += a[i] * b[i];
    double v[1] = {acc};
    block_partial<1, false>(v, partials);
}

// ------------------------------------------------------- finishing kernel

enum FinMode {
    FIN_GAMMA_INIT = 0,  // gamma = r'z at CG set-up
    FIN_GAMMA_ITER = 1,  // gamma_next = r'z after an update (stopping test, beta)
    FIN_PAP = 2,         // p'Ap: zero-curvature test, alpha
    FIN_LZ_GAMMA0 = 3,   // Lanczos: gamma_0 = r'r
    FIN_LZ_ALPHA = 4,    // Lanczos: alpha_k = gamma / p'(D^-1 A)p
    FIN_LZ_BETA = 5,     // Lanczos: beta_k = r'r / gamma
    FIN_PICARD = 6,      // Picard: max|h - h_prev|, max|h| -> converged flag
    FIN_IR_FIRST = 7,    // refinement: initial FP64 residual norm
    FIN_IR = 8,          // refinement: FP64 residual norm -> done flag
    FIN_NN_SCALE = 9,    // learned: image scale s = sqrt(sum r^2/aP / k)
    FIN_NN_LS = 10,      // learned: gamma = e'r / e'Ae (0 when not a descent step)
    FIN_FCG_BETA = 11,   // flexible CG (truncation 1): beta = -z'Ap / p'Ap
    FIN_GAMMA_NN = 12,   // gamma = r'z after the learned preconditioner (no convergence test)
    FIN_NN_LS_PC = 13,   // learned preconditioner: gamma_pc = e'r / e'Ae (0 when e'Ae <= 0)
    FIN_NN_SCALE_PC = 14,  // learned preconditioner: image scale of r
    FIN_GERSH = 15,      // max of the Gershgorin row bounds of D^-1 A
};

// The nv (1 or 2) totals of nb per-workgroup partials (value i at
// partials[i * nb + block]), summed or maximised in a fixed order by one
// (64, 4) workgroup; valid in thread 0.
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partials, int nb, int nv, bool is_max,
                                                double (&acc)[2]) {
    const int tid = threadIdx.x + threadIdx.y * WAVE;
    acc[0] = 0.0;
    acc[1] = 0.0;
    // each thread takes j = tid, tid + 256, ... in that order; the loads of 16
    // consecutive terms are issued together (the order of the sum is unchanged)
    constexpr int UF = 16, ST = WAVE * WAVES;
    for (int i = 0; i < nv; ++i) {
        double a = 0.0;
        bool first = true;
        for (int j0 = tid; j0 < nb; j0 += UF * ST) {
            double v[UF];
#pragma unroll
            for (int u = 0; u < UF; ++u) v[u] = partials[(size_t)i * nb + min(j0 + u * ST, nb - 1)];
#pragma unroll
            for (int u = 0; u < UF; ++u) {
                if (j0 + u * ST >= nb) break;
                const double b = v[u];
                if (first) { a = b; first = false; }
                else a = is_max ? ((isnan(b) || b > a) ? b : a) : a + b;
            }
        }
        acc[i] = a;
    }
    if (is_max) block_reduce<2, true>(acc);
    else if (nv == 2) block_reduce<2, false>(acc);
    else {
        double one[1] = {acc[0]};
        block_reduce<1, false>(one);
        acc[0] = one[0];
    }
}

// Two norms from the partials of sums of squares (k_true_resid):
// out[0] = sqrt(total 0), out[1] = sqrt(total 1).
// The Picard flag from max|h - h_prev| (v0) and max|h| (v1)
// (src/elliptic_solver.jl:153-169).
__device__ __forceinline__ void picard_flag(Scalars* sc, double v0, double v1) {
    sc->dmax = v0;
    sc->hmax = v1;
    sc->picard_nan = isnan(v0) || isnan(v1);
    sc->picard_converged = (!isnan(v0) && !isnan(v1) && v0 / (v1 + EPS64) < sc->picard_tol) ? 1 : 0;
}

// ppart: also the Picard test's maxima of k_true_resid, into sc.
__global__ __launch_bounds__(256) void k_finish_norms(const double* __restrict__ partials, int nb,
                                                      double* __restrict__ out, const double* __restrict__ ppart,
                                                      Scalars* sc) {
    double acc[2];
    reduce_partials(partials, nb, 2, false, acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        out[0] = sqrt(acc[0]);
        out[1] = sqrt(acc[1]);
    }
    if (ppart) {
        double m[2];
        reduce_partials(ppart, nb, 2, true, m);
        if (threadIdx.x == 0 && threadIdx.y == 0) picard_flag(sc, m[0], m[1]);
    }
}

// The per-solve reset of the device scalars (the fields a solve starts
// from), written on the device so that a solve's start needs no host round
// trip.  One thread.
__global__ void k_reset_scalars(Scalars* sc, double atol, double rtol, int itmax) {
    sc->done = 0;
    sc->solved = 0;
    sc->zero_curv = 0;
    sc->tired = 0;
    sc->iter = 0;
    sc->atol = atol;
    sc->rtol = rtol;
    sc->itmax = itmax;
    sc->counter = 0;
    sc->gpend = 0;
    sc->done_defer = 0;
    sc->breakdown = 0;
    sc->xpend_n = 0;
}

// One workgroup: sums (or maximises) the nb partials of the preceding
// reduction kernel in a fixed order and updates the solver scalars.
// Modes that follow an iteration kernel do nothing once the solve is done
// (that kernel returned early and wrote no partials).
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ partials, int nb, int mode, int k,
                                                Scalars* sc) {
    if ((mode == FIN_GAMMA_ITER || mode == FIN_PAP || mode == FIN_FCG_BETA || mode == FIN_GAMMA_NN ||
         mode == FIN_NN_LS_PC || mode == FIN_NN_SCALE_PC) &&
        sc->done)
        return;
    const bool is_max = mode == FIN_PICARD || mode == FIN_GERSH;
    const int nv = (mode == FIN_PICARD || mode == FIN_NN_LS || mode == FIN_NN_LS_PC) ? 2 : 1;
    double acc[2];
    reduce_partials(partials, nb, nv, is_max, acc);
    if (threadIdx.x != 0 || threadIdx.y != 0) return;
    const double v0 = acc[0], v1 = acc[1];
    switch (mode) {
        case FIN_GAMMA_INIT: fin_gamma(sc, v0, 0); break;
        case FIN_GAMMA_ITER: fin_gamma(sc, v0, 1); break;
        case FIN_PAP:
            sc->pAp = v0;
            if (!isfinite(v0)) {
                sc->breakdown = 2;
                sc->breakdown_value = v0;
                sc->done = 1;
            } else if (v0 <= EPS64 * sc->pNorm2 && fabs(v0) <= EPS64 * sc->pNorm2) {
                sc->zero_curv = 1;
                sc->done = 1;
            } else {
                sc->alpha = sc->gamma / v0;
            }
            break;
        case FIN_LZ_GAMMA0:
            sc->lz_gamma = v0;
            sc->beta = 0.0;
            break;
        case FIN_LZ_ALPHA: sc->lz_alpha[k] = sc->lz_gamma / v0; break;
        case FIN_LZ_BETA: {
            const double b = v0 / sc->lz_gamma;
            sc->lz_beta[k] = b;
            sc->beta = b;
            sc->lz_gamma = v0;
            break;
        }
        case FIN_IR_FIRST:
        case FIN_IR: {
            const double rn = sqrt(v0);
            sc->ir_rnorm = rn;
            if (mode == FIN_IR_FIRST) sc->ir_rnorm0 = rn;
            sc->ir_done = (rn <= sc->ir_atol + sc->ir_rtol * sc->ir_rnorm0) || (rn + 1.0 <= 1.0);
            break;
        }
        case FIN_NN_SCALE: sc->nn_scale = sqrt(v0 / (double)k); break;
        case FIN_FCG_BETA: sc->beta = -v0 / sc->pAp; break;
        case FIN_NN_SCALE_PC: sc->nn_scale_pc = sqrt(v0 / (double)k); break;
        case FIN_NN_LS_PC: {
            // any finite step along e lowers the energy by (e'r)^2 / e'Ae
            const double g = v0 / v1;
            sc->nn_gamma_pc = (v1 > 0.0 && isfinite(g)) ? g : 0.0;
            break;
        }
        case FIN_GAMMA_NN: {
            // r'z of the composite preconditioner: alpha = r'z / p'Ap is still
            // the exact line search along p (r is orthogonal to the previous p),
            // whatever its sign; the convergence test waits for the next
            // iteration, whose norm is the fixed preconditioner's
            const double beta = v0 / sc->gamma;
            sc->pNorm2 = fabs(v0) + beta * beta * sc->pNorm2;
            sc->gamma = v0;
            sc->beta = beta;
            sc->iter += 1;
            sc->tired = sc->iter >= sc->itmax;
            sc->done = sc->tired;
            break;
        }
        case FIN_NN_LS: {
            const double g = v0 / v1;
            const bool ok = v1 > 0.0 && v0 > 0.0 && isfinite(g);
            sc->nn_gamma = ok ? g : 0.0;
            sc->nn_fallback = ok ? 0 : 1;
            break;
        }
        case FIN_GERSH: sc->gersh = v0; break;
        case FIN_PICARD: picard_flag(sc, v0, v1); break;
        default: break;
    }
}

static inline void finish(const Geom& G, double* partials, int mode, int k, Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, G.nblocks, mode, k, sc);
}

// ------------------------------------------------------------ launchers

static inline dim3 cells_grid(int nx, int ny) { return dim3((nx + 63) / 64, (ny + 3) / 4); }

Geom make_geom(int nx, int ny, int R_override) {
    Geom G;
    G.nx = nx;
    G.ny = ny;
    int R;
    if (R_override > 0) R = R_override;
    else {
        // aim for >= 8192 waves (32 per CU on 256 CUs, enough loads in
        // flight to cover HBM latency), at most 32 rows per wave
        const long waves_per_row_strip = (nx + WAVE - 1) / WAVE;
        long r = ((long)ny * waves_per_row_strip) / 8192;
        R = (int)(r < 1 ? 1 : (r > 32 ? 32 : r));
        const char* env = getenv("SHK_MARCH_R");
        if (env && atoi(env) > 0) R = atoi(env);
    }
    G.R = R;
    G.block = dim3(WAVE, WAVES);
    G.grid = dim3((nx + WAVE - 1) / WAVE, (ny + WAVES * R - 1) / (WAVES * R));
    G.nblocks = G.grid.x * G.grid.y;
    return G;
}

// Geometry of the two-columns-per-lane FP64 stencil kernels: 128-column
// strips, R chosen for >= 8192 waves as in make_geom.
// Used where the arrays stream from HBM (stream_nt): inside the Infinity
// Cache the one-column kernels measured faster (6.4 against 5.8 TB/s for the
// operator at 4096 x 1024).
static bool wide_off() {
    static const bool off = getenv("SHK_NO_WIDE") != nullptr;
    return off;
}
bool wide_ok(const Geom& G) { return !wide_off() && G.nx % 2 == 0 && G.nx >= 2 && stream_nt(G); }
// Tiles of four 128-column strips by R rows (wide_tile); R for about 4096
// waves, at most 8 (at 4096 x 1024, R = 8 measured faster than 4 or 16).
Geom wide_geom(const Geom& G) {
    Geom W = G;
    const long strips = (G.nx + 2 * WAVE - 1) / (2 * WAVE);
    long r = ((long)G.ny * strips) / 4096;
    W.R = (int)(r < 1 ? 1 : (r > 8 ? 8 : r));
    W.grid = dim3((unsigned)((strips + WAVES - 1) / WAVES), (unsigned)((G.ny + W.R - 1) / W.R));
    W.nblocks = W.grid.x * W.grid.y;
    return W;
}
// The CG matvec (seven arrays per cell) also takes the wide kernel on large
// grids inside the Infinity Cache: 41 against 45 us at 4096 x 1024.
bool cg_wide_ok(const Geom& G) {
    return !wide_off() && G.nx % 2 == 0 && G.nx >= 2 && (stream_nt(G) || (long)G.nx * G.ny >= (1L << 21));
}
// Non-temporal streaming once the operator's arrays (40 B per cell) exceed
// the 256 MB Infinity Cache.
bool stream_nt(const Geom& G) {
    static const bool off = getenv("SHK_NO_NT") != nullptr;
    return !off && (double)G.nx * G.ny * 40.0 > 256.0 * 1024 * 1024;
}
// The CG matvec of FP64 CG with the V-cycle streams its coefficients and its
// p and Ap stores non-temporally once a CG iteration's arrays (about 72 B per
// cell: the FP64 coefficients, x, r, p, Ap, z and the finest V-cycle level)
// exceed the Infinity Cache: what the matvec streams then no longer evicts
// what the V-cycle reuses.  At 4096 x 1024 this took the form from 0.0793 to
// 0.0772 s per step (the finest down-leg 22.0 -> 20.4 us, the matvec
// 40.7 -> 39.8 us in the solve), although the matvec timed alone with the
// caches flushed is slightly slower with the hints (4.55 against 4.69 TB/s).
// (SHK_NO_NT is read at each call: a context created after it is set runs
// without the hints.)
static bool cg_stream_nt(const Geom& G) {
    return getenv("SHK_NO_NT") == nullptr && (double)G.nx * G.ny * 72.0 > 256.0 * 1024 * 1024;
}

// Geometry of the fused update+Chebyshev kernel: strips of 64-2H owned
// columns, R owned rows per wave.
Geom fused_geom(const Geom& G, int H) {
    Geom F = G;
    const int stride = WAVE - 2 * H;
    const int gx = (G.nx + stride - 1) / stride;
    // each wave recomputes 2H rows of warm-up/drain around its R owned rows
    int R = 2 * H + 2;
    if (R > 32) R = 32;
    const char* env = getenv("SHK_FUSED_R");
    if (env && atoi(env) > 0) R = atoi(env);
    F.R = R;
    F.grid = dim3(gx, (G.ny + WAVES * R - 1) / (WAVES * R));
    F.nblocks = F.grid.x * F.grid.y;
    return F;
}

namespace launch {

void ic_cells(const Phys& P, Fields F, hipStream_t st) {
    hipLaunchKernelGGL(k_ic_cells, cells_grid(P.nx, P.ny), dim3(64, 4), 0, st, P, F);
}
void stagger_b(const Phys& P, Fields F, hipStream_t st) {
    hipLaunchKernelGGL(k_stagger_b, cells_grid(P.nx + 1, P.ny + 1), dim3(64, 4), 0, st, P, F);
}
void stagger(int nx, int ny, const double* c, double* fx, double* fy, hipStream_t st) {
    hipLaunchKernelGGL(k_stagger, cells_grid(nx + 1, ny + 1), dim3(64, 4), 0, st, nx, ny, c, fx, fy);
}
void derived(const Geom& G, const Phys& P, Fields F, const double* h_prev, int mode, int check, double* partials,
             Scalars* sc, hipStream_t st) {
    const dim3 fgrid = cells_grid(G.nx + 1, G.ny + 1), cgrid = cells_grid(G.nx, G.ny), blk(WAVE, WAVES);
    if (mode == 0) hipLaunchKernelGGL((k_derived_faces<0>), fgrid, blk, 0, st, P, F);
    else hipLaunchKernelGGL((k_derived_faces<1>), fgrid, blk, 0, st, P, F);
    if (mode == 0 && check) {
        hipLaunchKernelGGL((k_derived_cells<0, true>), cgrid, blk, 0, st, P, F, h_prev, partials);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)(cgrid.x * cgrid.y),
                           (int)FIN_PICARD, 0, sc);
    } else if (mode == 0) {
        hipLaunchKernelGGL((k_derived_cells<0, false>), cgrid, blk, 0, st, P, F, h_prev, partials);
    } else {
        hipLaunchKernelGGL((k_derived_cells<1, false>), cgrid, blk, 0, st, P, F, h_prev, partials);
    }
}
void derived_march(const Phys& P, Fields F, int mode, bool full, hipStream_t st) {
    static const int r_env = [] {
        const char* e = getenv("SHK_DERIVED_R");
        return e ? atoi(e) : 0;
    }();
    const int R = r_env > 0 ? r_env : 8;
    const dim3 grid((P.nx + DER_OWN - 1) / DER_OWN, ((P.ny + R - 1) / R + WAVES - 1) / WAVES), blk(WAVE, WAVES);
    // one instantiation per sliding law (the initial-condition refresh, once, keeps the switch)
    auto go = [&](auto law) {
        constexpr int L = decltype(law)::value;
        if (full) hipLaunchKernelGGL((k_derived_march<0, true, L>), grid, blk, 0, st, P, F, R);
        else hipLaunchKernelGGL((k_derived_march<0, false, L>), grid, blk, 0, st, P, F, R);
    };
    if (mode == 1) {
        hipLaunchKernelGGL((k_derived_march<1, true, -1>), grid, blk, 0, st, P, F, R);
        return;
    }
    switch (P.sliding_law) {
        case SHK_SLIDING_REGULARIZED_COULOMB: go(std::integral_constant<int, SHK_SLIDING_REGULARIZED_COULOMB>{}); break;
        case SHK_SLIDING_PRESCRIBED: go(std::integral_constant<int, SHK_SLIDING_PRESCRIBED>{}); break;
        case SHK_SLIDING_LINEAR: go(std::integral_constant<int, SHK_SLIDING_LINEAR>{}); break;
        case SHK_SLIDING_REGULARIZED_COULOMB_FIELD:
            go(std::integral_constant<int, SHK_SLIDING_REGULARIZED_COULOMB_FIELD>{});
            break;
        case SHK_SLIDING_REGULARIZED_COULOMB_V0: go(std::integral_constant<int, SHK_SLIDING_REGULARIZED_COULOMB_V0>{}); break;
        default: go(std::integral_constant<int, -1>{}); break;
    }
}
void picard_check(const Phys& P, const double* h, const double* h_prev, double* partials, Scalars* sc,
                  hipStream_t st) {
    const dim3 cgrid = cells_grid(P.nx, P.ny);
    hipLaunchKernelGGL(k_picard_check, cgrid, dim3(WAVE, WAVES), 0, st, P.nx, P.ny, h, h_prev, partials);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)(cgrid.x * cgrid.y),
                       (int)FIN_PICARD, 0, sc);
}
void gap(const Phys& P, Fields F, double dt, hipStream_t st) {
    const int n = P.nx * P.ny;
    hipLaunchKernelGGL(k_gap, dim3((n + 255) / 256), dim3(256), 0, st, P, F, dt);
}
void fill_ieb(int n, double* ieb, double u, const double* moulin, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_ieb, dim3((n + 255) / 256), dim3(256), 0, st, n, ieb, u, moulin);
}
void to_f32(size_t n, const double* a, float* b, hipStream_t st) {
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b);
}
void ir_residual(const Geom& G, const double* aP, const double* aE, const double* aN, const double* b,
                 const double* x, float* r32, int first, double* partials, Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_ir_residual, G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, b, x, r32, partials);
    finish(G, partials, first ? FIN_IR_FIRST : FIN_IR, 0, sc, st);
}
void ir_correct(size_t n, double* x, const float* d, hipStream_t st) {
    hipLaunchKernelGGL(k_ir_correct, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, x, d);
}
void relax(size_t n, double* h, const double* hp, double a, hipStream_t st) {
    hipLaunchKernelGGL(k_relax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, h, hp, a);
}
void copy(const double* a, double* b, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b);
}
void zero(double* p, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p);
}

template <class T>
void assemble(const Geom& G, const Phys& P, Fields F, int parabolic, double dt, T* aP, T* aE, T* aN, T* rhs,
              hipStream_t st, double* cP, const AsmLow& lo, const AsmInit& in) {
    const dim3 grid = cells_grid(G.nx, G.ny), blk(WAVE, WAVES);
    AsmLow l = lo;
    if (!cP) l = AsmLow{};  // the copies come from the flux form's c
    const AsmInit i = sizeof(T) == 8 ? in : AsmInit{};
    if (parabolic)
        hipLaunchKernelGGL((k_assemble<true, T>), grid, blk, 0, st, P, F, dt, aP, aE, aN, rhs, cP, l, i);
    else
        hipLaunchKernelGGL((k_assemble<false, T>), grid, blk, 0, st, P, F, dt, aP, aE, aN, rhs, cP, l, i);
}
template <class T>
void matvec(const Geom& G, const T* aP, const T* aE, const T* aN, const T* x, T* y, hipStream_t st) {
    if constexpr (sizeof(T) == 8) {
        if (wide_ok(G)) {
            const Geom W = wide_geom(G);
            if (stream_nt(G))
                hipLaunchKernelGGL(k_matvec2<true>, W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, x, y);
            else
                hipLaunchKernelGGL(k_matvec2<false>, W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, x, y);
            return;
        }
    }
    hipLaunchKernelGGL(k_matvec<T>, G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, x, y);
}
template <class T>
void lanczos_init(const Geom& G, const T* rhs, const T* dg, int sym, T* r, T* p0, double* partials, Scalars* sc,
                  hipStream_t st) {
    if (sym)
        hipLaunchKernelGGL((k_lz_init<T, true>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, rhs, dg, r, p0, partials,
                           sc);
    else
        hipLaunchKernelGGL((k_lz_init<T, false>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, rhs, dg, r, p0, partials,
                           sc);
    finish(G, partials, FIN_LZ_GAMMA0, 0, sc, st);
}
template <class T>
void lanczos_step_a(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, int sym, const T* r,
                    const T* p_old, T* p_new, T* Ap, int k, double* partials, Scalars* sc, hipStream_t st) {
    if (sym)
        hipLaunchKernelGGL((k_lz_step_a<T, true>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, dg, r, p_old,
                           p_new, Ap, k, partials, sc);
    else
        hipLaunchKernelGGL((k_lz_step_a<T, false>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, dg, r,
                           p_old, p_new, Ap, k, partials, sc);
    finish(G, partials, FIN_LZ_ALPHA, k, sc, st);
}
template <class T>
void lanczos_step_b(const Geom& G, const T* Ap, const T* dg, int sym, T* r, int k, double* partials, Scalars* sc,
                    hipStream_t st) {
    if (sym)
        hipLaunchKernelGGL((k_lz_step_b<T, true>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, Ap, dg, r, k, partials,
                           sc);
    else
        hipLaunchKernelGGL((k_lz_step_b<T, false>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, Ap, dg, r, k, partials,
                           sc);
    finish(G, partials, FIN_LZ_BETA, k, sc, st);
}
void cheb_bounds(int nsteps, int degree, int use_estimate, int mode, Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_cheb_bounds, dim3(1), dim3(64), 0, st, nsteps, degree, use_estimate, mode, sc);
}
template <class T>
void gershgorin(const Geom& G, const T* dg, const T* aE, const T* aN, double* partials, Scalars* sc,
                hipStream_t st) {
    const dim3 grid = cells_grid(G.nx, G.ny);
    hipLaunchKernelGGL(k_gersh<T>, grid, dim3(WAVE, WAVES), 0, st, G.nx, G.ny, dg, aE, aN, partials);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)(grid.x * grid.y),
                       (int)FIN_GERSH, 0, sc);
}
int residual_norms_nblocks(const Geom& G) { return cg_wide_ok(G) ? wide_geom(G).nblocks : G.nblocks; }
void residual_norms(const Geom& G, const double* aP, const double* aE, const double* aN, const double* b,
                    const double* x, const double* x0, double* partials, double* out, hipStream_t st, double* h_out,
                    double relax, double* ppart, Scalars* sc) {
    if (!h_out || !sc) ppart = nullptr;
    int nb = G.nblocks;
    if (cg_wide_ok(G)) {
        const Geom W = wide_geom(G);
        hipLaunchKernelGGL(k_true_resid2, W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, b, x, x0, partials,
                           h_out, relax, ppart);
        nb = W.nblocks;
    } else {
        hipLaunchKernelGGL(k_true_resid, G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, b, x, x0, partials,
                           h_out, relax, ppart);
    }
    hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(WAVE, WAVES), 0, st, partials, nb, out, ppart, sc);
}
void reset_scalars(Scalars* sc, double atol, double rtol, int itmax, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(1), 0, st, sc, atol, rtol, itmax);
}
template <class T>
void cg_init(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, const T* b, const T* x0, T* r, T* z,
             T* xacc, T* p_zero, int jacobi, double* partials, Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_init<T>, G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, dg, b, x0, r, z, xacc,
                       p_zero, jacobi, partials, sc);
    if (jacobi) finish(G, partials, FIN_GAMMA_INIT, 0, sc, st);
}
template <class T>
void cheb_pass(const Geom& G, const T* aP, const T* aE, const T* aN, const T* dg, const T* r, const T* s_in,
               const T* pc_in, T* s_out, T* pc_out, T* y, int k, int degree, int mode, double* partials, Scalars* sc,
               hipStream_t st) {
    hipLaunchKernelGGL(k_cheb_pass<T>, G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, dg, r, s_in, pc_in,
                       s_out, pc_out, y, k, degree, mode, partials, sc);
    if (k == degree - 1 && mode >= 0) finish(G, partials, mode == 0 ? FIN_GAMMA_INIT : FIN_GAMMA_ITER, 0, sc, st);
}
template <class T>
void cheb_degree1(const Geom& G, const T* dg, const T* r, T* y, int mode, double* partials, Scalars* sc,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_cheb_deg1<T>, G.grid, G.block, 0, st, G.nx, G.ny, G.R, dg, r, y, mode, partials, sc);
    if (mode >= 0) finish(G, partials, mode == 0 ? FIN_GAMMA_INIT : FIN_GAMMA_ITER, 0, sc, st);
}
template <class T>
void cg_matvec(const Geom& G, const T* aP, const T* aE, const T* aN, const T* z, const T* p_old, T* p_new, T* Ap,
               double* partials, Scalars* sc, hipStream_t st) {
    if constexpr (sizeof(T) == 8) {
        if (cg_wide_ok(G)) {
            const Geom W = wide_geom(G);
            if (stream_nt(G))
                hipLaunchKernelGGL((k_cg_matvec2<true, false>), W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN,
                                   z, p_old, p_new, Ap, partials, sc, nullptr, 0, nullptr);
            else
                hipLaunchKernelGGL((k_cg_matvec2<false, false>), W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE,
                                   aN, z, p_old, p_new, Ap, partials, sc, nullptr, 0, nullptr);
            finish(W, partials, FIN_PAP, 0, sc, st);
            return;
        }
    }
    hipLaunchKernelGGL((k_cg_matvec<T, false>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, z, p_old, p_new,
                       Ap, partials, sc, nullptr, 0, nullptr);
    finish(G, partials, FIN_PAP, 0, sc, st);
}
template <class T>
void cg_update(const Geom& G, const T* dg, const T* p, const T* Ap, T* x, T* r, T* z, int jacobi, double* partials,
               Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL((k_cg_update<T, false>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, dg, p, Ap, x, r, z, jacobi,
                       partials, sc, nullptr, 0, nullptr, 0);
    if (jacobi) finish(G, partials, FIN_GAMMA_ITER, 0, sc, st);
}
int cg_matvec_deferred(const Geom& G, const double* aP, const double* aE, const double* aN, const double* z,
                       const double* p_old, double* p_new, double* Ap, double* pb, Scalars* sc, const double* pa,
                       int nba, double* x, hipStream_t st) {
    if (cg_wide_ok(G)) {
        const Geom W = wide_geom(G);
        if (cg_stream_nt(G))
            hipLaunchKernelGGL((k_cg_matvec2<true, true>), W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, z,
                               p_old, p_new, Ap, pb, sc, pa, nba, x);
        else
            hipLaunchKernelGGL((k_cg_matvec2<false, true>), W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, z,
                               p_old, p_new, Ap, pb, sc, pa, nba, x);
        return W.nblocks;
    }
    hipLaunchKernelGGL((k_cg_matvec<double, true>), G.grid, G.block, 0, st, G.nx, G.ny, G.R, aP, aE, aN, z, p_old,
                       p_new, Ap, pb, sc, pa, nba, x);
    return G.nblocks;
}
// The same with the V-cycle's FP32 z (cg_wide_ok grids only); xs as
// k_cg_matvec2's XS.
int cg_matvec_deferred(const Geom& G, const double* aP, const double* aE, const double* aN, const float* z,
                       const double* p_old, double* p_new, double* Ap, double* pb, Scalars* sc, const double* pa,
                       int nba, double* x, hipStream_t st, int xs) {
    const Geom W = wide_geom(G);
#define SHK_MV2(NT, XS)                                                                                          \
    hipLaunchKernelGGL((k_cg_matvec2<NT, true, float, XS>), W.grid, W.block, 0, st, W.nx, W.ny, W.R, aP, aE, aN, z, \
                       p_old, p_new, Ap, pb, sc, pa, nba, x)
    const bool nt = cg_stream_nt(G);
    if (xs == 0) {
        if (nt) SHK_MV2(true, 0);
        else SHK_MV2(false, 0);
    } else if (xs == 2) {
        if (nt) SHK_MV2(true, 2);
        else SHK_MV2(false, 2);
    } else {
        if (nt) SHK_MV2(true, 1);
        else SHK_MV2(false, 1);
    }
#undef SHK_MV2
    return W.nblocks;
}
void cg_defer_stop(size_t n, Scalars* sc, const double* pa, int nba, const double* p_old, double* x,
                   hipStream_t st, int xs, const double* p_pend) {
    const unsigned nblk = (unsigned)std::min<size_t>(1024, (n + 511) / 512);
    hipLaunchKernelGGL(k_cg_defer_stop, dim3(nblk), dim3(WAVE, WAVES), 0, st, n, sc, pa, nba, p_old, x, xs, p_pend);
}
void cg_xflush(size_t n, const Scalars* sc, double* x, hipStream_t st) {
    const unsigned nblk = (unsigned)std::min<size_t>(1024, (n + 511) / 512);
    hipLaunchKernelGGL(k_cg_xflush, dim3(nblk), dim3(WAVE, WAVES), 0, st, n, sc, x);
}
void cg_update_deferred(const Geom& G, const double* Ap, double* r, Scalars* sc, const double* pa, int nba,
                        const double* pb, int nbb, hipStream_t st) {
    // 1024 workgroups (one per 4 waves of the chip): each re-sums the partials
    // first, so more workgroups cost more than their better streaming gains
    // (in the solve at 4096 x 1024: 20.3 us at 1024, 20.7 at 2048, 23.1 at 4096)
    const size_t n = (size_t)G.nx * G.ny;
    const unsigned nblk = (unsigned)std::min<size_t>(1024, (n + 511) / 512);
    hipLaunchKernelGGL(k_cg_update_flat, dim3(nblk), dim3(WAVE, WAVES), 0, st, n, Ap, r, sc, pa, nba, pb, nbb);
}
template <class T>
int cg_update_cheb(const Geom& G, int degree, const T* aP, const T* aE, const T* aN, const T* dg, const T* p,
                   const T* Ap, const T* r_old, T* r_new, T* x, T* z, double* partials, Scalars* sc,
                   hipStream_t st) {
    const int H = degree - 1;
    const Geom F = fused_geom(G, H);
#define SHK_FUSED(HH)                                                                                             \
    case HH:                                                                                                      \
        hipLaunchKernelGGL((k_cg_update_cheb<T, HH>), F.grid, F.block, 0, st, G.nx, G.ny, F.R, aP, aE, aN, dg, p, \
                           Ap, r_old, r_new, x, z, partials, sc);                                                 \
        break;
    switch (H) {
        SHK_FUSED(1)
        SHK_FUSED(2)
        SHK_FUSED(3)
        SHK_FUSED(4)
        SHK_FUSED(5)
        SHK_FUSED(6)
        SHK_FUSED(7)
        default: return -1;
    }
#undef SHK_FUSED
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, F.nblocks, (int)FIN_GAMMA_ITER, 0, sc);
    return 0;
}

void mg_coarsen(int nxf, int nyf, int nxc, int nyc, const float* cf, const float* ef, const float* nf, float* cc,
                float* ec, float* nc, hipStream_t st) {
    hipLaunchKernelGGL(k_mg_coarsen, cells_grid(nxc, nyc), dim3(64, 4), 0, st, nxf, nyf, nxc, nyc, cf, ef, nf, cc,
                       ec, nc);
}
void mg_diag(int nx, int ny, const float* c, const float* e, const float* n, float* D, hipStream_t st) {
    hipLaunchKernelGGL(k_mg_diag, cells_grid(nx, ny), dim3(64, 4), 0, st, nx, ny, c, e, n, D);
}
void mg_coarsen_lin(int nxf, int nyf, int nxc, int nyc, const float* cf, const float* ef, const float* nf, float* cc,
                    float* ec, float* nc, hipStream_t st) {
    hipLaunchKernelGGL(k_mg_coarsen_lin, cells_grid(nxc, nyc), dim3(64, 4), 0, st, nxf, nyf, nxc, nyc, cf, ef, nf,
                       cc, ec, nc);
}
template <class TB, class CT>
void mg_restrict_lin(int nx, int ny, int nxc, int nyc, const CT* c, const CT* e, const CT* n, const TB* b,
                     const float* x, float* bc, const Scalars* sc, hipStream_t st) {
    hipLaunchKernelGGL((k_mg_restrict_lin<TB, CT>), cells_grid(nxc, nyc), dim3(64, 4), 0, st, nx, ny, nxc, nyc, c, e,
                       n, b, x, bc, sc);
}
template void mg_restrict_lin<float, float>(int, int, int, int, const float*, const float*, const float*,
                                            const float*, const float*, float*, const Scalars*, hipStream_t);
template void mg_restrict_lin<double, float>(int, int, int, int, const float*, const float*, const float*,
                                             const double*, const float*, float*, const Scalars*, hipStream_t);
template void mg_restrict_lin<float, uint16_t>(int, int, int, int, const uint16_t*, const uint16_t*,
                                               const uint16_t*, const float*, const float*, float*, const Scalars*,
                                               hipStream_t);
template void mg_restrict_lin<double, uint16_t>(int, int, int, int, const uint16_t*, const uint16_t*,
                                                const uint16_t*, const double*, const float*, float*,
                                                const Scalars*, hipStream_t);
void mg_prolong_lin(int nx, int ny, int nxc, float scale, float* x, const float* xc, const Scalars* sc,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_mg_prolong_lin, cells_grid(nx, ny), dim3(64, 4), 0, st, nx, ny, nxc, scale, x, xc, sc);
}
// SHK_MG_LIN_UNFUSED=1: the linear transfers always through their separate
// kernels (A/B measurements and the fused kernels' parity check).
static bool getenv_flag_lin_unfused() {
    static const bool v = getenv("SHK_MG_LIN_UNFUSED") != nullptr;
    return v;
}
// Largest level (cells) that runs the single-phase kernels; SHK_MG_FLAT
// overrides MG_FLAT_CELLS for A/B measurements.
static long mg_flat_cells() {
    static const long v = [] {
        const char* env = getenv("SHK_MG_FLAT");
        return env ? atol(env) : MG_FLAT_CELLS;
    }();
    return v;
}
// Levels above MG_FLAT_CELLS run the marching kernels on the CG geometry
// with R even.
static Geom mg_march_geom(int nx, int ny) {
    const Geom G = make_geom(nx, ny);
    return G.R % 2 == 0 ? G : make_geom(nx, ny, G.R + 1);
}
// Marching levels of even width take the two-column legs: tiles as the wide
// CG matvec, R even and chosen for about 8192 waves (at 4096 x 1024, in the
// solve: 30.6 / 28.6 us down / up, against 34.8 / 30.3 for 4096 waves and
// 32.5 / 30.0 for the one-column legs).  Same results as the one-column legs;
// also faster where the one-column grid leaves a partial round of
// workgroups (4160 x 1024: 37.5 against 45 us for the down-leg, caches cold).
static bool mg_wide(int nx) { return !wide_off() && nx % 2 == 0; }
static Geom mg_wide_geom(int nx, int ny) {
    Geom W = make_geom(nx, ny);
    const long strips = (nx + 2 * WAVE - 1) / (2 * WAVE);
    long r = ((long)ny * strips) / 8192;
    W.R = (int)(r < 2 ? 2 : (r > 8 ? 8 : r + (r & 1)));
    W.grid = dim3((unsigned)((strips + WAVES - 1) / WAVES), (unsigned)((ny + W.R - 1) / W.R));
    W.nblocks = W.grid.x * W.grid.y;
    return W;
}
// Geometry of the red-black two-column legs: strips of RB_OWN owned columns,
// R even, chosen for about 4096 waves and at most 8 rows.  In the solve at
// 4096 x 1024 the finest legs took 30.2 / 29.8 us (down / up) this way,
// against 36.5 / 36.2 for 8192 waves and 44.8 / 38.4 for 2048 waves of 16
// rows: the two rows of lookahead a wave reads before its first residual
// cost less over more rows.
static Geom mg_rb_geom(int nx, int ny, bool lin = false, int own_cols = 0) {
    Geom W = make_geom(nx, ny);
    const int own = own_cols ? own_cols : (lin ? RB_OWN_LIN : RB_OWN);
    const long strips = (nx + own - 1) / own;
    long r = ((long)ny * strips) / 4096;
    W.R = (int)(r < 2 ? 2 : (r > 8 ? 8 : r + (r & 1)));
    W.grid = dim3((unsigned)((strips + WAVES - 1) / WAVES), (unsigned)((ny + W.R - 1) / W.R));
    W.nblocks = W.grid.x * W.grid.y;
    return W;
}
// Red-black smoothing: the two-column marching legs on levels of even width
// above MG_FLAT_CELLS cells, the flat legs elsewhere (small or odd-width levels).
static bool mg_rb_flat(int nx, int ny) { return (long)nx * ny <= mg_flat_cells() || nx % 2 != 0; }
// The two-sweep finest legs: 120 owned columns per wave and R rows (even, 2
// to 16) for about 2048 waves, half mg_rb_geom's count: a wave's first eight
// marching steps only fill its five-stage pipeline, so longer waves pay off
// (4096 x 1024, fp64_mg s/step at R = 8 / 12 / 16 / 24 / 32: 0.0382 / 0.0367
// / 0.0360 / 0.0364 / 0.0383).  SHK_MG_SW2_R overrides R.
static Geom sw2_geom(int nx, int ny) {
    Geom W = mg_rb_geom(nx, ny, false, SW2_OWN);
    static const int r_env = [] {
        const char* e = getenv("SHK_MG_SW2_R");
        return e ? atoi(e) : 0;
    }();
    const long strips = (nx + SW2_OWN - 1) / SW2_OWN;
    const long r = ((long)ny * strips) / 2048;
    W.R = r_env >= 2 ? r_env + (r_env & 1) : (int)(r < 2 ? 2 : (r > 16 ? 16 : r + (r & 1)));
    W.grid.y = (unsigned)((ny + W.R - 1) / W.R);
    W.nblocks = W.grid.x * W.grid.y;
    return W;
}
// Workgroups (and so r'z partials) of the up-leg on an nx x ny level.
int mg_up_nblocks(int nx, int ny, int rb, int nu) {
    if (rb && nu == 2 && !mg_rb_flat(nx, ny)) return sw2_geom(nx, ny).nblocks;
    if (rb) return mg_rb_flat(nx, ny) ? (int)(cells_grid(nx, ny).x * cells_grid(nx, ny).y) : mg_rb_geom(nx, ny).nblocks;
    if ((long)nx * ny <= mg_flat_cells()) return (int)(cells_grid(nx, ny).x * cells_grid(nx, ny).y);
    return mg_wide(nx) ? mg_wide_geom(nx, ny).nblocks : mg_march_geom(nx, ny).nblocks;
}
// The finest level's red-black legs with BF16 coefficients (levels of even
// width that take the two-column legs; see mg_bf16_ok).
template <class TB>
void mg_down_bf16(int nx, int ny, int nxc, float omega, const uint16_t* c, const uint16_t* e, const uint16_t* n,
                  const uint16_t* D, const TB* b, float* x, float* bc, const Scalars* sc, int nu, hipStream_t st) {
    if (nu == 2) {
        const Geom W = sw2_geom(nx, ny);
        hipLaunchKernelGGL((k_mg_down_sw2<TB, uint16_t>), W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega, c, e, n,
                           D, b, x, bc, sc, CGUpd{});
        return;
    }
    const Geom W = mg_rb_geom(nx, ny);
    if (W.R == 2)
        hipLaunchKernelGGL((k_mg_down_rb2s<TB, uint16_t>), W.grid, W.block, 0, st, nx, ny, nxc, omega, c, e, n, D, b,
                           x, bc, sc);
    else
        hipLaunchKernelGGL((k_mg_down_rb2<TB, uint16_t>), W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega, c, e, n,
                           D, b, x, bc, sc, CGUpd{});
}
bool mg_down_fuses_update(int nx, int ny) { return mg_rb_geom(nx, ny).R > 2; }
void mg_down_bf16_upd(int nx, int ny, int nxc, float omega, const uint16_t* c, const uint16_t* e, const uint16_t* n,
                      const uint16_t* D, const double* b, float* x, float* bc, const CGUpd& U, int nu,
                      hipStream_t st) {
    if (nu == 2) {
        const Geom W = sw2_geom(nx, ny);
        hipLaunchKernelGGL((k_mg_down_sw2<double, uint16_t, true>), W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega,
                           c, e, n, D, b, x, bc, U.sc, U);
        return;
    }
    const Geom W = mg_rb_geom(nx, ny);  // R > 2 (mg_down_fuses_update)
    hipLaunchKernelGGL((k_mg_down_rb2<double, uint16_t, true>), W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega, c,
                       e, n, D, b, x, bc, U.sc, U);
}
template <class TB, class TO>
void mg_up_bf16(int nx, int ny, int nxc, float omega, float scale, const uint16_t* c, const uint16_t* e,
                const uint16_t* n, const uint16_t* D, const TB* b, const float* x, const float* xc, TO* out,
                int dot_mode, double* partials, Scalars* sc, int nu, hipStream_t st) {
    const Geom M = nu == 2 ? sw2_geom(nx, ny) : mg_rb_geom(nx, ny);
    const int with_dot = dot_mode == 2 ? 2 : (dot_mode >= 0 ? 1 : 0);
    if (nu == 2)
        hipLaunchKernelGGL((k_mg_up_sw2<TB, uint16_t, TO>), M.grid, M.block, 0, st, nx, ny, nxc, M.R, omega, scale, c,
                           e, n, D, b, x, xc, out, with_dot, partials, sc);
    else if (M.R == 2)
        hipLaunchKernelGGL((k_mg_up_rb2s<TB, uint16_t, TO>), M.grid, M.block, 0, st, nx, ny, nxc, omega, scale, c, e, n,
                           D, b, x, xc, out, with_dot, partials, sc);
    else
        hipLaunchKernelGGL((k_mg_up_rb2<TB, uint16_t, TO>), M.grid, M.block, 0, st, nx, ny, nxc, M.R, omega, scale, c,
                           e, n, D, b, x, xc, out, with_dot, partials, sc);
    if (dot_mode == 0 || dot_mode == 1)
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)M.nblocks,
                           dot_mode == 0 ? (int)FIN_GAMMA_INIT : (int)FIN_GAMMA_ITER, 0, sc);
}
template void mg_down_bf16<float>(int, int, int, float, const uint16_t*, const uint16_t*, const uint16_t*,
                                  const uint16_t*, const float*, float*, float*, const Scalars*, int, hipStream_t);
template void mg_down_bf16<double>(int, int, int, float, const uint16_t*, const uint16_t*, const uint16_t*,
                                   const uint16_t*, const double*, float*, float*, const Scalars*, int, hipStream_t);
template void mg_up_bf16<float, float>(int, int, int, float, float, const uint16_t*, const uint16_t*,
                                       const uint16_t*, const uint16_t*, const float*, const float*, const float*,
                                       float*, int, double*, Scalars*, int, hipStream_t);
template void mg_up_bf16<double, double>(int, int, int, float, float, const uint16_t*, const uint16_t*,
                                         const uint16_t*, const uint16_t*, const double*, const float*, const float*,
                                         double*, int, double*, Scalars*, int, hipStream_t);
template void mg_up_bf16<double, float>(int, int, int, float, float, const uint16_t*, const uint16_t*,
                                        const uint16_t*, const uint16_t*, const double*, const float*, const float*,
                                        float*, int, double*, Scalars*, int, hipStream_t);
// Levels whose red-black legs are the two-column kernels (and so can read BF16 coefficients).
bool mg_bf16_ok(int nx, int ny) { return !mg_rb_flat(nx, ny); }
// FP32 -> BF16, round to nearest even, four arrays of n values
__global__ __launch_bounds__(256) void k_to_bf16x4(size_t n, const float* __restrict__ a0, const float* __restrict__ a1,
                                                   const float* __restrict__ a2, const float* __restrict__ a3,
                                                   uint16_t* __restrict__ o0, uint16_t* __restrict__ o1,
                                                   uint16_t* __restrict__ o2, uint16_t* __restrict__ o3) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    auto rn = [](float v) -> uint16_t { return bf16_rne(v); };
    o0[i] = rn(a0[i]);
    o1[i] = rn(a1[i]);
    o2[i] = rn(a2[i]);
    o3[i] = rn(a3[i]);
}
void to_bf16x4(size_t n, const float* a0, const float* a1, const float* a2, const float* a3, uint16_t* o0,
               uint16_t* o1, uint16_t* o2, uint16_t* o3, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16x4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a0, a1, a2, a3, o0, o1,
                       o2, o3);
}
template <class TB>
void mg_down(int nx, int ny, int nxc, int nyc, float omega, const float* c, const float* e, const float* n,
             const float* D, const TB* b, float* x, float* bc, const Scalars* sc, int rb, hipStream_t st) {
    if (rb) {
        if (mg_rb_flat(nx, ny)) {
            hipLaunchKernelGGL(k_mg_down_rb_flat<TB>, cells_grid(nxc, nyc), dim3(WAVE, WAVES), 0, st, nx, ny, nxc,
                               nyc, omega, c, e, n, D, b, x, bc, sc);
        } else {
            const Geom W = mg_rb_geom(nx, ny);
            if (W.R == 2)
                hipLaunchKernelGGL((k_mg_down_rb2s<TB, float>), W.grid, W.block, 0, st, nx, ny, nxc, omega, c, e, n, D, b, x,
                                   bc, sc);
            else
                hipLaunchKernelGGL((k_mg_down_rb2<TB, float>), W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega, c, e, n, D, b,
                                   x, bc, sc, CGUpd{});
        }
        return;
    }
    if ((long)nx * ny <= mg_flat_cells()) {
        hipLaunchKernelGGL(k_mg_down_flat<TB>, cells_grid(nxc, nyc), dim3(WAVE, WAVES), 0, st, nx, ny, nxc, nyc,
                           omega, c, e, n, D, b, x, bc, sc);
        return;
    }
    if (mg_wide(nx)) {
        const Geom W = mg_wide_geom(nx, ny);
        hipLaunchKernelGGL(k_mg_down_march2<TB>, W.grid, W.block, 0, st, nx, ny, nxc, W.R, omega, c, e, n, D, b, x,
                           bc, sc);
        return;
    }
    const Geom M = mg_march_geom(nx, ny);
    hipLaunchKernelGGL(k_mg_down_march<TB>, M.grid, M.block, 0, st, nx, ny, nxc, M.R, omega, c, e, n, D, b, x, bc,
                       sc);
}
template <class TB>
void mg_up(int nx, int ny, int nxc, float omega, float scale, const float* c, const float* e, const float* n,
           const float* D, const TB* b, const float* x, const float* xc, TB* out, int dot_mode,
           double* partials, Scalars* sc, int rb, hipStream_t st) {
    const bool flat = rb ? mg_rb_flat(nx, ny) : (long)nx * ny <= mg_flat_cells();
    const bool wide = !flat && (rb || mg_wide(nx));
    const Geom M = flat ? Geom{} : (rb ? mg_rb_geom(nx, ny) : (wide ? mg_wide_geom(nx, ny) : mg_march_geom(nx, ny)));
    const dim3 grid = flat ? cells_grid(nx, ny) : M.grid;
    // dot_mode 2: the partials of r'z stay pending for the next CG matvec
    const int with_dot = dot_mode == 2 ? 2 : (dot_mode >= 0 ? 1 : 0);
    if (rb && flat)
        hipLaunchKernelGGL(k_mg_up_rb_flat<TB>, grid, dim3(WAVE, WAVES), 0, st, nx, ny, nxc, omega, scale, c, e, n,
                           D, b, x, xc, out, with_dot, partials, sc);
    else if (rb && M.R == 2)
        hipLaunchKernelGGL((k_mg_up_rb2s<TB, float, TB>), grid, M.block, 0, st, nx, ny, nxc, omega, scale, c, e, n, D, b, x, xc,
                           out, with_dot, partials, sc);
    else if (rb)
        hipLaunchKernelGGL((k_mg_up_rb2<TB, float, TB>), grid, M.block, 0, st, nx, ny, nxc, M.R, omega, scale, c, e, n, D, b, x,
                           xc, out, with_dot, partials, sc);
    else if (flat)
        hipLaunchKernelGGL(k_mg_up_flat<TB>, grid, dim3(WAVE, WAVES), 0, st, nx, ny, nxc, omega, scale, c, e, n, D,
                           b, x, xc, out, with_dot, partials, sc);
    else if (wide)
        hipLaunchKernelGGL(k_mg_up_march2<TB>, grid, M.block, 0, st, nx, ny, nxc, M.R, omega, scale, c, e, n, D, b,
                           x, xc, out, with_dot, partials, sc);
    else
        hipLaunchKernelGGL(k_mg_up_march<TB>, grid, M.block, 0, st, nx, ny, nxc, M.R, omega, scale, c, e, n, D, b,
                           x, xc, out, with_dot, partials, sc);
    if (dot_mode == 0 || dot_mode == 1)
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)(grid.x * grid.y),
                           dot_mode == 0 ? (int)FIN_GAMMA_INIT : (int)FIN_GAMMA_ITER, 0, sc);
}
// The red-black legs of a level that reaches the next one by the linear
// transfers, where they are fused into the two-row kernels (even width above
// MG_FLAT_CELLS cells, two rows per wave); false where they are not, and the
// caller takes the plain legs with the separate transfer kernels.
// nu = 2: two sweeps each way (k_mg_down_rbn / k_mg_up_rbn) where the level
// takes two rows per wave with their narrower owned strip; one otherwise.
bool mg_lin_nu2(int nx, int ny) {
    return !mg_rb_flat(nx, ny) && !getenv_flag_lin_unfused() && mg_rb_geom(nx, ny, true, RBN<2>::OWN).R == 2;
}
// Rows per wave of the two-sweep legs: four on levels of more than
// MG_RBN4_CELLS cells, where the legs are throughput-bound and four rows share
// the halo rows two would re-read; two below, where they are latency-bound.
static int mg_rbn_rows(int nx, int ny) {
    static const long cells = [] {
        const char* env = getenv("SHK_MG_RBN4");
        return env ? atol(env) : MG_RBN4_CELLS;
    }();
    return (long)nx * ny > cells ? 4 : 2;
}
static Geom mg_rbn_geom(int nx, int ny) {
    Geom W = mg_rb_geom(nx, ny, true, RBN<2>::OWN);
    W.R = mg_rbn_rows(nx, ny);
    W.grid.y = (unsigned)((ny + W.R - 1) / W.R);
    W.nblocks = W.grid.x * W.grid.y;
    return W;
}
template <class TB, class CT>
bool mg_down_lin(int nx, int ny, int nxc, float omega, const CT* c, const CT* e, const CT* n, const CT* D,
                 const TB* b, float* x, float* bc, const Scalars* sc, int nu, hipStream_t st) {
    if constexpr (std::is_same<CT, float>::value) {
        if (nu == 2 && mg_lin_nu2(nx, ny)) {
            const Geom W = mg_rbn_geom(nx, ny);
            if (W.R == 4)
                hipLaunchKernelGGL((k_mg_down_rbn<TB, 2, 4>), W.grid, W.block, 0, st, nx, ny, nxc, omega, c, e, n, D,
                                   b, x, bc, sc);
            else
                hipLaunchKernelGGL((k_mg_down_rbn<TB, 2, 2>), W.grid, W.block, 0, st, nx, ny, nxc, omega, c, e, n, D,
                                   b, x, bc, sc);
            return true;
        }
    }
    if (mg_rb_flat(nx, ny) || getenv_flag_lin_unfused()) return false;
    const Geom W = mg_rb_geom(nx, ny, true);
    if (W.R != 2) return false;
    hipLaunchKernelGGL((k_mg_down_rb2s<TB, CT, true>), W.grid, W.block, 0, st, nx, ny, nxc, omega, c, e, n, D, b, x,
                       bc, sc);
    return true;
}
template <class TB, class CT, class TO>
bool mg_up_lin(int nx, int ny, int nxc, float omega, float scale, const CT* c, const CT* e, const CT* n,
               const CT* D, const TB* b, const float* x, const float* xc, TO* out, int dot_mode, double* partials,
               Scalars* sc, int nu, hipStream_t st) {
    const bool nu2 = std::is_same<CT, float>::value && nu == 2 && mg_lin_nu2(nx, ny);
    if (!nu2 && (mg_rb_flat(nx, ny) || getenv_flag_lin_unfused())) return false;
    const Geom M = nu2 ? mg_rbn_geom(nx, ny) : mg_rb_geom(nx, ny, true);
    if (!nu2 && M.R != 2) return false;
    const int with_dot = dot_mode == 2 ? 2 : (dot_mode >= 0 ? 1 : 0);
    if constexpr (std::is_same<CT, float>::value) {
        if (nu2 && M.R == 4)
            hipLaunchKernelGGL((k_mg_up_rbn<TB, TO, 2, 4>), M.grid, M.block, 0, st, nx, ny, nxc, omega, scale, c, e,
                               n, D, b, x, xc, out, with_dot, partials, sc);
        else if (nu2)
            hipLaunchKernelGGL((k_mg_up_rbn<TB, TO, 2, 2>), M.grid, M.block, 0, st, nx, ny, nxc, omega, scale, c, e,
                               n, D, b, x, xc, out, with_dot, partials, sc);
        else
            hipLaunchKernelGGL((k_mg_up_rb2s<TB, CT, TO, true>), M.grid, M.block, 0, st, nx, ny, nxc, omega, scale,
                               c, e, n, D, b, x, xc, out, with_dot, partials, sc);
    } else {
        hipLaunchKernelGGL((k_mg_up_rb2s<TB, CT, TO, true>), M.grid, M.block, 0, st, nx, ny, nxc, omega, scale, c, e,
                           n, D, b, x, xc, out, with_dot, partials, sc);
    }
    if (dot_mode == 0 || dot_mode == 1)
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)M.nblocks,
                           dot_mode == 0 ? (int)FIN_GAMMA_INIT : (int)FIN_GAMMA_ITER, 0, sc);
    return true;
}
template bool mg_down_lin<float, float>(int, int, int, float, const float*, const float*, const float*, const float*,
                                        const float*, float*, float*, const Scalars*, int, hipStream_t);
template bool mg_up_lin<float, float, float>(int, int, int, float, float, const float*, const float*, const float*,
                                             const float*, const float*, const float*, const float*, float*, int,
                                             double*, Scalars*, int, hipStream_t);
template void mg_down<float>(int, int, int, int, float, const float*, const float*, const float*, const float*,
                             const float*, float*, float*, const Scalars*, int, hipStream_t);
template void mg_down<double>(int, int, int, int, float, const float*, const float*, const float*, const float*,
                              const double*, float*, float*, const Scalars*, int, hipStream_t);
template void mg_up<float>(int, int, int, float, float, const float*, const float*, const float*, const float*,
                           const float*, const float*, const float*, float*, int, double*, Scalars*, int,
                           hipStream_t);
template void mg_up<double>(int, int, int, float, float, const float*, const float*, const float*, const float*,
                            const double*, const float*, const float*, double*, int, double*, Scalars*, int,
                            hipStream_t);
size_t mg_tail_stride(size_t tail_cells) { return (tail_cells + 3) & ~(size_t)3; }
// the packed block (c, e, n, D, inverse) in floats, rounded up to 16 B
static size_t mg_tail_pack(size_t tail_cells, size_t coarsest_cells) {
    return (4 * mg_tail_stride(tail_cells) + coarsest_cells * coarsest_cells + 3) & ~(size_t)3;
}
size_t mg_tail_lds(size_t tail_cells, size_t coarsest_cells, size_t hcells) {
    // rhs, x, the packed block and the restriction's staging rows
    return (2 * mg_tail_stride(tail_cells) + mg_tail_pack(tail_cells, coarsest_cells) + hcells) * sizeof(float);
}
int mg_tail_reserve(size_t bytes) {
    return hipFuncSetAttribute((const void*)k_mg_tail, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
                   hipSuccess
               ? 0
               : -1;
}
void mg_tail(const MGTail& T, const Scalars* sc, hipStream_t st) {
    size_t cells = 0;
    for (int l = 0; l < T.nlev; ++l) cells += (size_t)T.L[l].nx * T.L[l].ny;
    const size_t nc = (size_t)T.L[T.nlev - 1].nx * T.L[T.nlev - 1].ny;
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(1024), mg_tail_lds(cells, nc, (size_t)T.hcells), st, T, sc);
}
void mg_dense_inverse(int nx, int ny, const float* c, const float* e, const float* n, float* inv, hipStream_t st) {
    hipLaunchKernelGGL(k_mg_dense_inverse_lds, dim3(1), dim3(WAVE, 16), 0, st, nx, ny, c, e, n, inv);
}

static inline unsigned nn_blocks(size_t n) {
    size_t b = (n + WAVE * WAVES - 1) / (WAVE * WAVES);
    return (unsigned)(b > 1024 ? 1024 : (b < 1 ? 1 : b));
}
void nn_resid_scale(size_t n, const double* aP, const double* rhs, const double* Ap, double* r, double* partials,
                    Scalars* sc, hipStream_t st) {
    const unsigned nb = nn_blocks(n);
    hipLaunchKernelGGL(k_nn_resid, dim3(nb), dim3(WAVE, WAVES), 0, st, n, aP, rhs, Ap, r, partials);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)nb, (int)FIN_NN_SCALE, (int)n, sc);
}
void nn_features(int nx, int ny, int W8, int H8, const double* aP, const double* aE, const double* aN,
                 const double* cP, const double* r, const Scalars* sc, int slot, void* x, hipStream_t st) {
    const size_t np = (size_t)W8 * H8;
    hipLaunchKernelGGL(k_nn_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, nx, ny, W8, H8, aP, aE,
                       aN, cP, r, sc, slot, (nn_bf16x8*)x);
}
void nn_correction(int nx, int ny, int W8, const float* N, const double* aP, const double* aE, const double* aN,
                   const Scalars* sc, int slot, double* e, hipStream_t st) {
    const size_t n = (size_t)nx * ny;
    hipLaunchKernelGGL(k_nn_correction, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nx, ny, W8, N, aP, aE,
                       aN, sc, slot, e);
}
void nn_line_search(size_t n, const double* e, const double* r, const double* Ae, double* partials, Scalars* sc,
                    int slot, hipStream_t st) {
    const unsigned nb = nn_blocks(n);
    hipLaunchKernelGGL(k_nn_ls, dim3(nb), dim3(WAVE, WAVES), 0, st, n, e, r, Ae, partials);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)nb,
                       (int)(slot ? FIN_NN_LS_PC : FIN_NN_LS), 0, sc);
}
void nn_warm(size_t n, const double* h, const double* e, const Scalars* sc, double* x0, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_warm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, h, e, sc, x0);
}
void nn_target(int nx, int ny, int W8, int H8, const double* aP, const double* h_new, const double* h_prev,
               const Scalars* sc, float* t, hipStream_t st) {
    const size_t np = (size_t)W8 * H8;
    hipLaunchKernelGGL(k_nn_target, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, nx, ny, W8, H8, aP, h_new,
                       h_prev, sc, t);
}
void nn_target_cg(int nx, int ny, int W8, int H8, const double* aP, const double* x_final, const double* x_k,
                  const double* s_k, float* t, hipStream_t st) {
    const size_t np = (size_t)W8 * H8;
    hipLaunchKernelGGL(k_nn_target_cg, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, nx, ny, W8, H8, aP,
                       x_final, x_k, s_k, t);
}
void nn_rscale(size_t n, const double* aP, const double* r, double* partials, Scalars* sc, hipStream_t st) {
    const unsigned nb = nn_blocks(n);
    hipLaunchKernelGGL(k_nn_rscale, dim3(nb), dim3(WAVE, WAVES), 0, st, n, aP, r, partials, sc);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)nb, (int)FIN_NN_SCALE_PC, (int)n,
                       sc);
}
void nn_r1(size_t n, const double* r, const double* Ae, const Scalars* sc, double* r1, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_r1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, r, Ae, sc, r1);
}
void nn_combine(size_t n, const double* r, const double* e, double* z, double* partials, Scalars* sc,
                hipStream_t st) {
    const unsigned nb = nn_blocks(n);
    hipLaunchKernelGGL(k_nn_combine, dim3(nb), dim3(WAVE, WAVES), 0, st, n, r, e, z, partials, sc);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)nb, (int)FIN_GAMMA_NN, 0, sc);
}
void fcg_beta(size_t n, const double* z, const double* Ap, double* partials, Scalars* sc, hipStream_t st) {
    const unsigned nb = nn_blocks(n);
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(WAVE, WAVES), 0, st, n, z, Ap, partials, sc);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(WAVE, WAVES), 0, st, partials, (int)nb, (int)FIN_FCG_BETA, 0, sc);
}
void nn_ref(Scalars* sc, int mode, double n, hipStream_t st) {
    hipLaunchKernelGGL(k_nn_ref, dim3(1), dim3(64), 0, st, sc, mode, n);
}

void read_flush(size_t n, const double* a, double* sink, hipStream_t st) {
    hipLaunchKernelGGL(k_read_flush, dim3(4096), dim3(256), 0, st, n, a, sink);
}
void cg_finish(int n, const double* x, const double* h_prev, double* h_new, double relax, hipStream_t st) {
    hipLaunchKernelGGL(k_cg_finish, dim3((n + 255) / 256), dim3(256), 0, st, n, x, h_prev, h_new, relax);
}

#define SHK_INST(T)                                                                                               \
    template void assemble<T>(const Geom&, const Phys&, Fields, int, double, T*, T*, T*, T*, hipStream_t,         \
                              double*, const AsmLow&, const AsmInit&);                                            \
    template void matvec<T>(const Geom&, const T*, const T*, const T*, const T*, T*, hipStream_t);                \
    template void lanczos_init<T>(const Geom&, const T*, const T*, int, T*, T*, double*, Scalars*, hipStream_t);  \
    template void lanczos_step_a<T>(const Geom&, const T*, const T*, const T*, const T*, int, const T*, const T*, \
                                    T*, T*, int, double*, Scalars*, hipStream_t);                                 \
    template void lanczos_step_b<T>(const Geom&, const T*, const T*, int, T*, int, double*, Scalars*,             \
                                    hipStream_t);                                                                 \
    template void gershgorin<T>(const Geom&, const T*, const T*, const T*, double*, Scalars*, hipStream_t);       \
    template void cg_init<T>(const Geom&, const T*, const T*, const T*, const T*, const T*, const T*, T*, T*, T*,  \
                             T*, int, double*, Scalars*, hipStream_t);                                            \
    template void cheb_pass<T>(const Geom&, const T*, const T*, const T*, const T*, const T*, const T*, const T*, \
                               T*, T*, T*, int, int, int, double*, Scalars*, hipStream_t);                        \
    template void cheb_degree1<T>(const Geom&, const T*, const T*, T*, int, double*, Scalars*, hipStream_t);      \
    template void cg_matvec<T>(const Geom&, const T*, const T*, const T*, const T*, const T*, T*, T*, double*,    \
                               Scalars*, hipStream_t);                                                            \
    template void cg_update<T>(const Geom&, const T*, const T*, const T*, T*, T*, T*, int, double*, Scalars*,     \
                               hipStream_t);                                                                      \
    template int cg_update_cheb<T>(const Geom&, int, const T*, const T*, const T*, const T*, const T*, const T*,  \
                                   const T*, T*, T*, T*, double*, Scalars*, hipStream_t);
SHK_INST(double)
SHK_INST(float)
#undef SHK_INST

}  // namespace launch
}  // namespace shk
