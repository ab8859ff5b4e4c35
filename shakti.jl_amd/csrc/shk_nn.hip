// The learned part: a 4-level U-Net (16/32/64/128 channels) for gfx950 with
// hand-written BF16 MFMA kernels (FP32 accumulation) for inference and
// training.  No reference counterpart; the solver-side hooks are in
// shk_api.hip and the feature / warm-start kernels in shk_kernels.hip.
//
// Images are NHWC BF16: element (b, y, x, c) at ((b*H + y)*W + x)*C + c.
//
// 3x3 convolution as an implicit GEMM on v_mfma_f32_16x16x32_bf16.  The GEMM
// is D[co][pixel] = sum_k Wg[co][k] * X[k][pixel] with k = tap*CIN + ci
// (tap = 3*(dy+1) + (dx+1)), padded to KPAD = a multiple of 32 with zero
// weights.  Lane l of a wave supplies, per 32-wide k chunk:
//   A (weights):  row co = 16*cb + l%16, k = 8*(l/16) .. +7  -> one 16 B load
//   B (image):    col pixel = l%16,      k = 8*(l/16) .. +7  -> 8 channels of
//                 one pixel at one tap, one 16 B load (CIN % 8 == 0)
// and receives D rows co = 16*cb + 4*(l/16) + i (i = 0..3) of pixel l%16, so
// each lane stores 4 consecutive channels of one pixel (8 B).
// The backward-data pass is the same kernel on the gradient with the weights
// transposed and the taps flipped; the weight gradient contracts over pixels
// (A = dY^T, B = shifted X) with per-workgroup partial sums reduced in a fixed
// order, so every result is deterministic.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shakti_learned.h"
#include "shk_internal.hpp"

namespace shknn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct LayerDef {
    int ca, cb, cout, up, act, level;
};
// encoder a/b pairs per level, bottleneck, decoder pairs (first input
// upsampled from the level below and concatenated with the skip), 1x1 head
constexpr LayerDef LAYERS[SHK_L_COUNT] = {
    {8, 0, 16, 0, 1, 0},     {16, 0, 16, 0, 1, 0},  {16, 0, 32, 0, 1, 1},  {32, 0, 32, 0, 1, 1},
    {32, 0, 64, 0, 1, 2},    {64, 0, 64, 0, 1, 2},  {64, 0, 128, 0, 1, 3}, {128, 0, 128, 0, 1, 3},
    {128, 64, 64, 1, 1, 2},  {64, 0, 64, 0, 1, 2},  {64, 32, 32, 1, 1, 1}, {32, 0, 32, 0, 1, 1},
    {32, 16, 16, 1, 1, 0},   {16, 0, 16, 0, 1, 0},  {16, 0, 1, 0, 0, 0},
};

constexpr int kpad(int cin) { return ((9 * cin / 8 + 3) / 4) * 32; }
constexpr int64_t layer_params(int l) {
    return l == SHK_L_HEAD ? 17
                           : (int64_t)LAYERS[l].cout * 9 * (LAYERS[l].ca + LAYERS[l].cb) + LAYERS[l].cout;
}
constexpr int64_t layer_offset(int l) { return l == 0 ? 0 : layer_offset(l - 1) + layer_params(l - 1); }
constexpr int64_t NUM_PARAMS = layer_offset(SHK_L_COUNT);

struct TensorDef {
    int ch, level, f32;
};
constexpr TensorDef TENSORS[SHK_T_COUNT] = {
    {8, 0, 0},   {16, 0, 0},  {16, 0, 0},  {16, 1, 0},  {32, 1, 0},  {32, 1, 0},  {32, 2, 0},  {64, 2, 0},
    {64, 2, 0},  {64, 3, 0},  {128, 3, 0}, {128, 3, 0}, {64, 2, 0},  {64, 2, 0},  {32, 1, 0},  {32, 1, 0},
    {16, 0, 0},  {16, 0, 0},  {1, 0, 1},   // OUT
    {1, 0, 1},   {16, 0, 0},  {16, 0, 0},  {48, 0, 0},  {32, 1, 0},  {32, 1, 0},  {96, 1, 0},  {64, 2, 0},
    {64, 2, 0},  {192, 2, 0}, {128, 3, 0}, {128, 3, 0}, {64, 3, 0},  {64, 2, 0},  {64, 2, 0},  {32, 2, 0},
    {32, 1, 0},  {32, 1, 0},  {16, 1, 0},  {16, 0, 0},  {16, 0, 0},
};

// forward input / output tensors and backward gradient tensors of each layer
struct LayerIO {
    int in_a, in_b, out, d_out, d_in;
};
constexpr LayerIO IO[SHK_L_COUNT] = {
    {SHK_T_X, -1, SHK_T_A0, SHK_T_DA0, -1},
    {SHK_T_A0, -1, SHK_T_S0, SHK_T_DS0, SHK_T_DA0},
    {SHK_T_P0, -1, SHK_T_A1, SHK_T_DA1, SHK_T_DP0},
    {SHK_T_A1, -1, SHK_T_S1, SHK_T_DS1, SHK_T_DA1},
    {SHK_T_P1, -1, SHK_T_A2, SHK_T_DA2, SHK_T_DP1},
    {SHK_T_A2, -1, SHK_T_S2, SHK_T_DS2, SHK_T_DA2},
    {SHK_T_P2, -1, SHK_T_A3, SHK_T_DA3, SHK_T_DP2},
    {SHK_T_A3, -1, SHK_T_T3, SHK_T_DT3, SHK_T_DA3},
    {SHK_T_T3, SHK_T_S2, SHK_T_U2, SHK_T_DU2, SHK_T_DCAT2},
    {SHK_T_U2, -1, SHK_T_V2, SHK_T_DV2, SHK_T_DU2},
    {SHK_T_V2, SHK_T_S1, SHK_T_U1, SHK_T_DU1, SHK_T_DCAT1},
    {SHK_T_U1, -1, SHK_T_V1, SHK_T_DV1, SHK_T_DU1},
    {SHK_T_V1, SHK_T_S0, SHK_T_U0, SHK_T_DU0, SHK_T_DCAT0},
    {SHK_T_U0, -1, SHK_T_V0, SHK_T_DV0, SHK_T_DU0},
    {SHK_T_V0, -1, SHK_T_OUT, SHK_T_DOUT, SHK_T_DV0},
};

// ----------------------------------------------------------------- kernels

// y = act(conv3x3([up(xa), xb]) + bias).  CA channels from xa (read at
// (y/2, x/2) of a half-resolution image when UPA), CB from xb.  One wave per
// 16*PB pixels of one row, all COUT channels; 4 waves per workgroup.
template <int CA, int CB, int COUT, int PB, bool UPA, bool ACT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16* __restrict__ xa, const bf16* __restrict__ xb,
                                                 const bf16* __restrict__ wg, const float* __restrict__ bias,
                                                 bf16* __restrict__ y, int B, int H, int W, float slope) {
    constexpr int CIN = CA + CB;
    static_assert(CA % 8 == 0 && CB % 8 == 0 && COUT % 16 == 0, "channel counts");
    constexpr int GPT = CIN / 8;  // 8-channel groups per tap
    constexpr int G = 9 * GPT;
    constexpr int NCH = (G + 3) / 4;
    constexpr int KP = NCH * 32;
    constexpr int NCB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_x = (W + 16 * PB - 1) / (16 * PB);
    const int row = wave / tiles_x;
    if (row >= B * H) return;
    const int tx = wave - row * tiles_x;
    const int b = row / H, yc = row - b * H;
    const int x0 = tx * 16 * PB;
    const int px = lane & 15, q = lane >> 4;
    const int Hh = H >> 1, Wh = W >> 1;
    f32x4 acc[PB][NCB];
#pragma unroll
    for (int i = 0; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < NCB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    for (int c = 0; c < NCH; ++c) {
        const int g = 4 * c + q;
        const bool gv = g < G;
        const int tap = gv ? g / GPT : 0;
        const int ci0 = (g - (g / GPT) * GPT) * 8;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        bf16x8 wf[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
            wf[cb] = *(const bf16x8*)(wg + (size_t)(cb * 16 + px) * KP + (size_t)g * 8);
        const int yy = yc + dy;
        const bool yv = gv && yy >= 0 && yy < H;
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
            const int xx = x0 + pb * 16 + px + dx;
            bf16x8 v = zero8;
            if (yv && xx >= 0 && xx < W) {
                if (CB == 0 || ci0 < CA) {
                    if (UPA)
                        v = *(const bf16x8*)(xa + (((size_t)b * Hh + (yy >> 1)) * Wh + (xx >> 1)) * CA + ci0);
                    else
                        v = *(const bf16x8*)(xa + (((size_t)b * H + yy) * W + xx) * CA + ci0);
                } else {
                    v = *(const bf16x8*)(xb + (((size_t)b * H + yy) * W + xx) * CB + (ci0 - CA));
                }
            }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                acc[pb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[cb], v, acc[pb][cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int pb = 0; pb < PB; ++pb) {
        const int xx = x0 + pb * 16 + px;
        if (xx >= W) continue;
        bf16* dst = y + (((size_t)b * H + yc) * W + xx) * COUT + 4 * q;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            bf16x4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[pb][cb][i];
                if (bias) v += bias[cb * 16 + 4 * q + i];
                if (ACT) v = v > 0.f ? v : slope * v;
                o[i] = (bf16)v;
            }
            *(bf16x4*)(dst + cb * 16) = o;
        }
    }
}

// Weight-gradient partials: part[s][co][col] = sum over the pixels of split s
// of dy[p][co] * X[p shifted by tap(col)][ci(col)], col = tap*CIN + ci.  One
// wave per workgroup computes a (16*MB) x 32 tile.  Needs W % 8 == 0 and
// (B*H*W) % 32 == 0 so that each lane's 8 pixels lie in one row.
template <int CA, int CB, int COUT, bool UPA>
__global__ __launch_bounds__(64) void k_conv3x3_dw(const bf16* __restrict__ xa, const bf16* __restrict__ xb,
                                                   const bf16* __restrict__ dy, float* __restrict__ part, int B,
                                                   int H, int W, long nps) {
    constexpr int CIN = CA + CB;
    constexpr int NCOL = 9 * CIN;
    constexpr int NCOLP = ((NCOL + 31) / 32) * 32;
    constexpr int MB = COUT >= 32 ? 2 : 1;
    const int lane = threadIdx.x;
    const int px = lane & 15, q = lane >> 4;
    const int col0 = blockIdx.x * 32, co0 = blockIdx.y * 16 * MB;
    const long P = (long)B * H * W;
    const long p_begin = (long)blockIdx.z * nps;
    const long p_end = p_begin + nps < P ? p_begin + nps : P;
    const int Hh = H >> 1, Wh = W >> 1;
    int tapn[2], cin[2];
    bool colv[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        const int col = col0 + nb * 16 + px;
        colv[nb] = col < NCOL;
        tapn[nb] = colv[nb] ? col / CIN : 0;
        cin[nb] = colv[nb] ? col - tapn[nb] * CIN : 0;
    }
    f32x4 acc[MB][2];
#pragma unroll
    for (int i = 0; i < MB; ++i) acc[i][0] = acc[i][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long p0 = p_begin; p0 < p_end; p0 += 32) {
        const long pg = p0 + 8 * q;
        const int b = (int)(pg / ((long)H * W));
        const long rem = pg - (long)b * H * W;
        const int yc = (int)(rem / W), xc = (int)(rem - (long)yc * W);
        bf16x8 af[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const bf16* src = dy + (size_t)pg * COUT + co0 + mb * 16 + px;
#pragma unroll
            for (int j = 0; j < 8; ++j) af[mb][j] = src[(size_t)j * COUT];
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            bf16x8 bv = {};
            const int ddy = tapn[nb] / 3 - 1, ddx = tapn[nb] % 3 - 1;
            const int yy = yc + ddy;
            if (colv[nb] && yy >= 0 && yy < H) {
                const int ci = cin[nb];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int xx = xc + j + ddx;
                    if (xx >= 0 && xx < W) {
                        if (CB == 0 || ci < CA) {
                            if (UPA)
                                bv[j] = xa[(((size_t)b * Hh + (yy >> 1)) * Wh + (xx >> 1)) * CA + ci];
                            else
                                bv[j] = xa[(((size_t)b * H + yy) * W + xx) * CA + ci];
                        } else {
                            bv[j] = xb[(((size_t)b * H + yy) * W + xx) * CB + (ci - CA)];
                        }
                    }
                }
            }
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], bv, acc[mb][nb], 0, 0, 0);
        }
    }
    float* dst = part + (size_t)blockIdx.z * COUT * NCOLP;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                dst[(size_t)(co0 + mb * 16 + 4 * q + i) * NCOLP + col0 + nb * 16 + px] = acc[mb][nb][i];
}

// grad[co*ncol + col] = sum_s part[s][co][col] (fixed order), col < ncol.
__global__ __launch_bounds__(256) void k_reduce_dw(const float* __restrict__ part, int S, int cout, int ncol,
                                                   int ncolp, float* __restrict__ grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cout * ncol) return;
    const int co = i / ncol, col = i - co * ncol;
    float a = 0.f;
    for (int s = 0; s < S; ++s) a += part[((size_t)s * cout + co) * ncolp + col];
    grad[i] = a;
}

// Bias-gradient partials: part[s][co] = sum over split s of dy[p][co].
template <int COUT>
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* __restrict__ dy, long P, long nps,
                                                   float* __restrict__ part) {
    constexpr int NG = COUT / 8;         // 8-channel groups
    constexpr int PL = 256 / NG;         // pixel lanes
    __shared__ float red[256][8];
    const int t = threadIdx.x;
    const int cg = t % NG, pl = t / NG;
    const long p_begin = (long)blockIdx.x * nps;
    const long p_end = p_begin + nps < P ? p_begin + nps : P;
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long p = p_begin + pl; p < p_end; p += PL) {
        const bf16x8 v = *(const bf16x8*)(dy + (size_t)p * COUT + cg * 8);
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] += (float)v[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[t][j] = a[j];
    __syncthreads();
    if (t < NG) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int l = 0; l < PL; ++l)
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += red[l * NG + t][j];
#pragma unroll
        for (int j = 0; j < 8; ++j) part[(size_t)blockIdx.x * COUT + t * 8 + j] = s[j];
    }
}

// 2x2 average pooling (C % 8 == 0): one thread per coarse pixel and 8 channels.
__global__ __launch_bounds__(256) void k_pool(const bf16* __restrict__ in, bf16* __restrict__ out, int B, int Hc,
                                              int Wc, int C) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int NG = C / 8;
    if (i >= (long)B * Hc * Wc * NG) return;
    const int g = (int)(i % NG);
    const long pc = i / NG;
    const int xc = (int)(pc % Wc);
    const long r = pc / Wc;
    const int yc = (int)(r % Hc), b = (int)(r / Hc);
    const int Wf = 2 * Wc, Hf = 2 * Hc;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const bf16x8 v =
                *(const bf16x8*)(in + (((size_t)b * Hf + 2 * yc + dy) * Wf + 2 * xc + dx) * C + g * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
        }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.25f * s[j]);
    *(bf16x8*)(out + (size_t)pc * C + g * 8) = o;
}

// Pooling backward, accumulated: dfine[p][c] += dcoarse[p/2][c] / 4.
__global__ __launch_bounds__(256) void k_pool_bwd_add(const bf16* __restrict__ dc, bf16* __restrict__ df, int B,
                                                      int Hf, int Wf, int C) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int NG = C / 8;
    if (i >= (long)B * Hf * Wf * NG) return;
    const int g = (int)(i % NG);
    const long pf = i / NG;
    const int xf = (int)(pf % Wf);
    const long r = pf / Wf;
    const int yf = (int)(r % Hf), b = (int)(r / Hf);
    const bf16x8 c = *(const bf16x8*)(dc + (((size_t)b * (Hf / 2) + yf / 2) * (Wf / 2) + xf / 2) * C + g * 8);
    bf16x8 f = *(const bf16x8*)(df + (size_t)pf * C + g * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) f[j] = (bf16)((float)f[j] + 0.25f * (float)c[j]);
    *(bf16x8*)(df + (size_t)pf * C + g * 8) = f;
}

// Backward of [nearest-upsample(deep), skip] concatenation: dcat [P_f][CA+CB]
// -> ddeep[p_c][c] = sum over the 2x2 fine pixels of dcat[.][c], c < CA, and
// dskip[p_f][c] = dcat[p_f][CA + c].  Thread per (coarse pixel, group) for
// the first part, per (fine pixel, group) for the second.
__global__ __launch_bounds__(256) void k_split_up(const bf16* __restrict__ dcat, bf16* __restrict__ ddeep, int B,
                                                  int Hc, int Wc, int CA, int CT) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int NG = CA / 8;
    if (i >= (long)B * Hc * Wc * NG) return;
    const int g = (int)(i % NG);
    const long pc = i / NG;
    const int xc = (int)(pc % Wc);
    const long r = pc / Wc;
    const int yc = (int)(r % Hc), b = (int)(r / Hc);
    const int Wf = 2 * Wc, Hf = 2 * Hc;
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const bf16x8 v =
                *(const bf16x8*)(dcat + (((size_t)b * Hf + 2 * yc + dy) * Wf + 2 * xc + dx) * CT + g * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
        }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (bf16)s[j];
    *(bf16x8*)(ddeep + (size_t)pc * CA + g * 8) = o;
}
__global__ __launch_bounds__(256) void k_split_skip(const bf16* __restrict__ dcat, bf16* __restrict__ dskip,
                                                    long P, int CA, int CB) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const int NG = CB / 8;
    if (i >= P * NG) return;
    const int g = (int)(i % NG);
    const long p = i / NG;
    *(bf16x8*)(dskip + (size_t)p * CB + g * 8) = *(const bf16x8*)(dcat + (size_t)p * (CA + CB) + CA + g * 8);
}

// LeakyReLU derivative applied in place: g *= (y > 0 ? 1 : slope).
__global__ __launch_bounds__(256) void k_act_grad(bf16* __restrict__ g, const bf16* __restrict__ y, long n8,
                                                  float slope) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n8) return;
    bf16x8 gv = *(const bf16x8*)(g + (size_t)i * 8);
    const bf16x8 yv = *(const bf16x8*)(y + (size_t)i * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) gv[j] = (bf16)((float)yv[j] > 0.f ? (float)gv[j] : slope * (float)gv[j]);
    *(bf16x8*)(g + (size_t)i * 8) = gv;
}

// 1x1 head: out[p] = bh + sum_c wh[c] v[p][c] (16 channels).
__global__ __launch_bounds__(256) void k_head(const bf16* __restrict__ v, const float* __restrict__ wh,
                                              float* __restrict__ out, long P) {
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const bf16x8 a = *(const bf16x8*)(v + (size_t)p * 16);
    const bf16x8 c = *(const bf16x8*)(v + (size_t)p * 16 + 8);
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += wh[j] * (float)a[j];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += wh[8 + j] * (float)c[j];
    out[p] = s + wh[16];
}

// Head backward: dv[p][c] = dout[p] wh[c]; partials (per workgroup) of
// sum_p dout[p] v[p][c] (16) and sum_p dout[p] (the bias).
__global__ __launch_bounds__(256) void k_head_bwd(const bf16* __restrict__ v, const float* __restrict__ wh,
                                                  const float* __restrict__ dout, bf16* __restrict__ dv, long P,
                                                  long nps, float* __restrict__ part) {
    __shared__ float red[256][17];
    const int t = threadIdx.x;
    const long p_begin = (long)blockIdx.x * nps;
    const long p_end = p_begin + nps < P ? p_begin + nps : P;
    float a[17];
#pragma unroll
    for (int j = 0; j < 17; ++j) a[j] = 0.f;
    for (long p = p_begin + t; p < p_end; p += 256) {
        const float d = dout[p];
        const bf16x8 v0 = *(const bf16x8*)(v + (size_t)p * 16);
        const bf16x8 v1 = *(const bf16x8*)(v + (size_t)p * 16 + 8);
        bf16x8 o0, o1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a[j] += d * (float)v0[j];
            a[8 + j] += d * (float)v1[j];
            o0[j] = (bf16)(d * wh[j]);
            o1[j] = (bf16)(d * wh[8 + j]);
        }
        a[16] += d;
        *(bf16x8*)(dv + (size_t)p * 16) = o0;
        *(bf16x8*)(dv + (size_t)p * 16 + 8) = o1;
    }
#pragma unroll
    for (int j = 0; j < 17; ++j) red[t][j] = a[j];
    __syncthreads();
    if (t < 17) {
        float s = 0.f;
        for (int l = 0; l < 256; ++l) s += red[l][t];
        part[(size_t)blockIdx.x * 17 + t] = s;
    }
}

// Per-image sums for the scale-free loss: part[b][blk][0..2] = sums of N t,
// N N, t t over the image's pixels in this workgroup's range.
__global__ __launch_bounds__(256) void k_loss_sums(const float* __restrict__ N, const float* __restrict__ t,
                                                   long HW, long nps, double* __restrict__ part) {
    __shared__ double red[256][3];
    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const long p_begin = (long)blockIdx.x * nps;
    const long p_end = p_begin + nps < HW ? p_begin + nps : HW;
    double s0 = 0, s1 = 0, s2 = 0;
    for (long p = p_begin + tid; p < p_end; p += 256) {
        const double n = N[(size_t)b * HW + p], tv = t[(size_t)b * HW + p];
        s0 += n * tv;
        s1 += n * n;
        s2 += tv * tv;
    }
    red[tid][0] = s0;
    red[tid][1] = s1;
    red[tid][2] = s2;
    __syncthreads();
    if (tid < 3) {
        double s = 0;
        for (int l = 0; l < 256; ++l) s += red[l][tid];
        part[((size_t)b * gridDim.x + blockIdx.x) * 3 + tid] = s;
    }
}

// Per image: L_b = 1 - c^2/(a s); gradient coefficients k1 = -2c/(a s B),
// k2 = 2c^2/(a^2 s B) so that dL/dN = k1 t + k2 N.  coef[b] = {L_b, k1, k2}.
__global__ void k_loss_finish(const double* __restrict__ part, int nblk, int B, double* __restrict__ coef) {
    const int b = threadIdx.x;
    if (b >= B) return;
    double c = 0, a = 0, s = 0;
    for (int k = 0; k < nblk; ++k) {
        c += part[((size_t)b * nblk + k) * 3 + 0];
        a += part[((size_t)b * nblk + k) * 3 + 1];
        s += part[((size_t)b * nblk + k) * 3 + 2];
    }
    const double tiny = 1e-300;
    const double as = a * s + tiny;
    coef[b * 3 + 0] = 1.0 - c * c / as;
    coef[b * 3 + 1] = -2.0 * c / (as * B);
    coef[b * 3 + 2] = 2.0 * c * c / ((a + tiny) * as * B);
}
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ N, const float* __restrict__ t,
                                                   const double* __restrict__ coef, long HW, int B,
                                                   float* __restrict__ dN) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= HW * B) return;
    const int b = (int)(i / HW);
    dN[i] = (float)(coef[b * 3 + 1] * (double)t[i] + coef[b * 3 + 2] * (double)N[i]);
}

// Energy-norm variant of the scale-free loss.  The input image carries the
// Jacobi-scaled operator: r^ = ch0, east / north couplings a = ch1 / 2, ch3 / 2,
// west / south ch2 / 2, ch4 / 2, and the reaction c = ch5.  In flux form the
// energy of u is a sum of non-negative terms,
//   E(u) = sum_cells c u^2 + sum_faces a_f (u_i - u_j)^2,
// so it stays accurate with BF16 couplings.  Per image:
//   L_b = 1 - (N.r^)^2 / (E(N) E(t)),
// the relative energy-norm error left after the best scaling of N when t
// solves A^ t = r^.  part[b][blk] = {N.r^, E(N), E(t)}.
__global__ __launch_bounds__(256) void k_eloss_sums(const float* __restrict__ N, const bf16* __restrict__ x,
                                                    const float* __restrict__ t, int H, int W, long nps,
                                                    double* __restrict__ part) {
    __shared__ double red[256][3];
    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const long HW = (long)H * W;
    const long p_begin = (long)blockIdx.x * nps;
    const long p_end = p_begin + nps < HW ? p_begin + nps : HW;
    const float* Nb = N + (size_t)b * HW;
    const float* tb = t + (size_t)b * HW;
    double s0 = 0, s1 = 0, s2 = 0;
    for (long p = p_begin + tid; p < p_end; p += 256) {
        const int yy = (int)(p / W), xx = (int)(p - (long)yy * W);
        const bf16x8 f = *(const bf16x8*)(x + ((size_t)b * HW + p) * 8);
        const double r = (float)f[0], aE = 0.5f * (float)f[1], aN = 0.5f * (float)f[3], c = (float)f[5];
        const double n = Nb[p], tv = tb[p];
        const double nE = xx + 1 < W ? Nb[p + 1] : 0.0, tE = xx + 1 < W ? tb[p + 1] : 0.0;
        const double nN = yy + 1 < H ? Nb[p + W] : 0.0, tN = yy + 1 < H ? tb[p + W] : 0.0;
        s0 += n * r;
        s1 += c * n * n + aE * (n - nE) * (n - nE) + aN * (n - nN) * (n - nN);
        s2 += c * tv * tv + aE * (tv - tE) * (tv - tE) + aN * (tv - tN) * (tv - tN);
    }
    red[tid][0] = s0;
    red[tid][1] = s1;
    red[tid][2] = s2;
    __syncthreads();
    if (tid < 3) {
        double s = 0;
        for (int l = 0; l < 256; ++l) s += red[l][tid];
        part[((size_t)b * gridDim.x + blockIdx.x) * 3 + tid] = s;
    }
}

// coef[b] = {L_b, k1, k2}: dL/dN = k1 r^ + k2 (A^ N), with A^ N in flux form.
__global__ void k_eloss_finish(const double* __restrict__ part, int nblk, int B, double* __restrict__ coef) {
    const int b = threadIdx.x;
    if (b >= B) return;
    double nr = 0, en = 0, et = 0;
    for (int k = 0; k < nblk; ++k) {
        nr += part[((size_t)b * nblk + k) * 3 + 0];
        en += part[((size_t)b * nblk + k) * 3 + 1];
        et += part[((size_t)b * nblk + k) * 3 + 2];
    }
    const double tiny = 1e-300;
    const double d = en * et + tiny;
    coef[b * 3 + 0] = 1.0 - nr * nr / d;
    coef[b * 3 + 1] = -2.0 * nr / (d * B);
    coef[b * 3 + 2] = 2.0 * nr * nr / ((en + tiny) * d * B);
}

__global__ __launch_bounds__(256) void k_eloss_grad(const float* __restrict__ N, const bf16* __restrict__ x,
                                                    const double* __restrict__ coef, int H, int W, int B,
                                                    float* __restrict__ dN) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long HW = (long)H * W;
    if (i >= HW * B) return;
    const int b = (int)(i / HW);
    const long p = i - (long)b * HW;
    const int yy = (int)(p / W), xx = (int)(p - (long)yy * W);
    const float* Nb = N + (size_t)b * HW;
    const bf16x8 f = *(const bf16x8*)(x + (size_t)i * 8);
    const double n = Nb[p];
    double an = (double)(float)f[5] * n;
    if (xx + 1 < W) an += 0.5 * (double)(float)f[1] * (n - Nb[p + 1]);
    if (xx > 0) an += 0.5 * (double)(float)f[2] * (n - Nb[p - 1]);
    if (yy + 1 < H) an += 0.5 * (double)(float)f[3] * (n - Nb[p + W]);
    if (yy > 0) an += 0.5 * (double)(float)f[4] * (n - Nb[p - W]);
    dN[i] = (float)(coef[b * 3 + 1] * (double)(float)f[0] + coef[b * 3 + 2] * an);
}

// Adam on the FP32 master weights (bias-corrected step sizes from the host).
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ th, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] = th[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 GEMM copies of one 3x3 layer's weights: wg[co][KP] (tap*CIN + ci,
// zero padded) and wt[ci][KPT] (flipped tap*COUT + co, for the
// backward-data pass).
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ w, int cin, int cout, int kp, int kpt,
                                              bf16* __restrict__ wg, bf16* __restrict__ wt) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long ng = (long)cout * kp, nt = (long)cin * kpt;
    if (i < ng) {
        const int co = (int)(i / kp), k = (int)(i - (long)co * kp);
        wg[i] = (bf16)(k < 9 * cin ? w[(size_t)co * 9 * cin + k] : 0.f);
    }
    if (wt && i < nt) {
        const int ci = (int)(i / kpt), k = (int)(i - (long)ci * kpt);
        float val = 0.f;
        if (k < 9 * cout) {
            const int tapf = k / cout, co = k - tapf * cout;
            val = w[(size_t)co * 9 * cin + (8 - tapf) * cin + ci];
        }
        wt[i] = (bf16)val;
    }
}

__global__ __launch_bounds__(256) void k_to_float(const bf16* __restrict__ a, float* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// ------------------------------------------------------------ dispatch

static inline unsigned blocks(long n, int per = 256) { return (unsigned)((n + per - 1) / per); }

template <int CA, int CB, int COUT, bool UPA, bool ACT>
void conv_fwd_t(const bf16* xa, const bf16* xb, const bf16* wg, const float* bias, bf16* y, int B, int H, int W,
                float slope, hipStream_t st) {
    constexpr int PB = COUT <= 48 ? 4 : (COUT <= 128 ? 2 : 1);
    const long waves = (long)B * H * ((W + 16 * PB - 1) / (16 * PB));
    hipLaunchKernelGGL((k_conv3x3<CA, CB, COUT, PB, UPA, ACT>), dim3(blocks(waves, 4)), dim3(256), 0, st, xa, xb,
                       wg, bias, y, B, H, W, slope);
}

template <int CA, int CB, int COUT, bool UPA>
void conv_dw_t(const bf16* xa, const bf16* xb, const bf16* dy, float* part, size_t part_cap, float* grad, int B,
               int H, int W, hipStream_t st) {
    constexpr int CIN = CA + CB;
    constexpr int NCOL = 9 * CIN;
    constexpr int NCOLP = ((NCOL + 31) / 32) * 32;
    constexpr int MB = COUT >= 32 ? 2 : 1;
    const int gx = NCOLP / 32, gy = COUT / (16 * MB);
    const long P = (long)B * H * W;
    long S = 2048 / (gx * gy);
    if (S < 1) S = 1;
    const long cap_s = (long)(part_cap / ((size_t)COUT * NCOLP));
    if (S > cap_s) S = cap_s;
    if (S > P / 32) S = P / 32;
    if (S < 1) S = 1;
    long nps = (P + S - 1) / S;
    nps = ((nps + 31) / 32) * 32;
    S = (P + nps - 1) / nps;
    hipLaunchKernelGGL((k_conv3x3_dw<CA, CB, COUT, UPA>), dim3(gx, gy, (unsigned)S), dim3(64), 0, st, xa, xb, dy,
                       part, B, H, W, nps);
    hipLaunchKernelGGL(k_reduce_dw, dim3(blocks((long)COUT * NCOL)), dim3(256), 0, st, part, (int)S, COUT, NCOL,
                       NCOLP, grad);
}

template <int COUT>
void bias_grad_t(const bf16* dy, long P, float* part, float* grad, hipStream_t st) {
    long S = P / 4096;
    if (S < 1) S = 1;
    if (S > 512) S = 512;
    const long nps = (P + S - 1) / S;
    hipLaunchKernelGGL(k_bias_grad<COUT>, dim3((unsigned)S), dim3(256), 0, st, dy, P, nps, part);
    hipLaunchKernelGGL(k_reduce_dw, dim3(blocks(COUT)), dim3(256), 0, st, part, (int)S, 1, COUT, COUT, grad);
}

void conv_fwd(int l, const bf16* xa, const bf16* xb, const bf16* wg, const float* bias, bf16* y, int B, int H,
              int W, float slope, hipStream_t st) {
    switch (l) {
        case SHK_L_ENC0A: conv_fwd_t<8, 0, 16, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC0B: conv_fwd_t<16, 0, 16, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC1A: conv_fwd_t<16, 0, 32, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC1B: conv_fwd_t<32, 0, 32, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC2A: conv_fwd_t<32, 0, 64, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC2B: conv_fwd_t<64, 0, 64, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC3A: conv_fwd_t<64, 0, 128, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_ENC3B: conv_fwd_t<128, 0, 128, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC2A: conv_fwd_t<128, 64, 64, true, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC2B: conv_fwd_t<64, 0, 64, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC1A: conv_fwd_t<64, 32, 32, true, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC1B: conv_fwd_t<32, 0, 32, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC0A: conv_fwd_t<32, 16, 16, true, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        case SHK_L_DEC0B: conv_fwd_t<16, 0, 16, false, true>(xa, xb, wg, bias, y, B, H, W, slope, st); break;
        default: break;
    }
}

// dX = conv3x3(dY, flipped transposed weights): input channels = the
// layer's COUT, output channels = its CIN (the concatenation for decoders).
void conv_bwd_data(int l, const bf16* dy, const bf16* wt, bf16* dx, int B, int H, int W, hipStream_t st) {
    switch (l) {
        case SHK_L_ENC0B: conv_fwd_t<16, 0, 16, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC1A: conv_fwd_t<32, 0, 16, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC1B: conv_fwd_t<32, 0, 32, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC2A: conv_fwd_t<64, 0, 32, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC2B: conv_fwd_t<64, 0, 64, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC3A: conv_fwd_t<128, 0, 64, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_ENC3B: conv_fwd_t<128, 0, 128, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC2A: conv_fwd_t<64, 0, 192, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC2B: conv_fwd_t<64, 0, 64, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC1A: conv_fwd_t<32, 0, 96, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC1B: conv_fwd_t<32, 0, 32, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC0A: conv_fwd_t<16, 0, 48, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        case SHK_L_DEC0B: conv_fwd_t<16, 0, 16, false, false>(dy, nullptr, wt, nullptr, dx, B, H, W, 0.f, st); break;
        default: break;
    }
}

void conv_dw(int l, const bf16* xa, const bf16* xb, const bf16* dy, float* part, size_t cap, float* grad, int B,
             int H, int W, hipStream_t st) {
    switch (l) {
        case SHK_L_ENC0A: conv_dw_t<8, 0, 16, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC0B: conv_dw_t<16, 0, 16, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC1A: conv_dw_t<16, 0, 32, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC1B: conv_dw_t<32, 0, 32, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC2A: conv_dw_t<32, 0, 64, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC2B: conv_dw_t<64, 0, 64, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC3A: conv_dw_t<64, 0, 128, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_ENC3B: conv_dw_t<128, 0, 128, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC2A: conv_dw_t<128, 64, 64, true>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC2B: conv_dw_t<64, 0, 64, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC1A: conv_dw_t<64, 32, 32, true>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC1B: conv_dw_t<32, 0, 32, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC0A: conv_dw_t<32, 16, 16, true>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        case SHK_L_DEC0B: conv_dw_t<16, 0, 16, false>(xa, xb, dy, part, cap, grad, B, H, W, st); break;
        default: break;
    }
}

void bias_grad(int cout, const bf16* dy, long P, float* part, float* grad, hipStream_t st) {
    switch (cout) {
        case 16: bias_grad_t<16>(dy, P, part, grad, st); break;
        case 32: bias_grad_t<32>(dy, P, part, grad, st); break;
        case 64: bias_grad_t<64>(dy, P, part, grad, st); break;
        case 128: bias_grad_t<128>(dy, P, part, grad, st); break;
        default: break;
    }
}

}  // namespace shknn

using namespace shknn;

// ------------------------------------------------------------ the network

struct shk_unet {
    shk_unet_cfg cfg{};
    hipStream_t own = nullptr;
    hipStream_t st = nullptr;
    std::string err;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    bf16* wg[SHK_L_COUNT] = {};
    bf16* wt[SHK_L_COUNT] = {};
    void* t[SHK_T_COUNT] = {};
    size_t t_elems[SHK_T_COUNT] = {};
    float* part = nullptr;
    size_t part_cap = 0;
    double* lpart = nullptr;
    double* lcoef = nullptr;
    size_t lpart_cap = 0;
    const void* x_in = nullptr;
    int B = 0, H = 0, W = 0;
    int adam_t = 0;
};

namespace {

int nn_fail(shk_unet* n, int code, const std::string& msg) {
    if (n) n->err = msg;
    return code;
}

#define NNCHK(call)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return nn_fail(net, SHK_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

inline const bf16* T(const shk_unet* n, int id) { return (const bf16*)n->t[id]; }
inline bf16* TW(shk_unet* n, int id) { return (bf16*)n->t[id]; }

int pack_weights(shk_unet* net) {
    for (int l = 0; l < SHK_L_HEAD; ++l) {
        const int cin = LAYERS[l].ca + LAYERS[l].cb, cout = LAYERS[l].cout;
        const int kp = kpad(cin), kpt = kpad(cout);
        const long nmax = std::max((long)cout * kp, (long)cin * kpt);
        hipLaunchKernelGGL(k_pack, dim3(blocks(nmax)), dim3(256), 0, net->st, net->theta + layer_offset(l), cin,
                           cout, kp, kpt, net->wg[l], net->wt[l]);
    }
    NNCHK(hipGetLastError());
    return SHK_OK;
}

int check_shape(shk_unet* net, int B, int H, int W) {
    if (B < 1 || B > net->cfg.max_batch || H < 32 || W < 64 || H > net->cfg.max_h || W > net->cfg.max_w ||
        H % 32 || W % 64)
        return nn_fail(net, SHK_ERR_ARG,
                       "image shape out of range: need 1 <= B <= max_batch, H a multiple of 32 and W a multiple "
                       "of 64 within the capacity");
    return SHK_OK;
}

}  // namespace

namespace shknn {

int forward(shk_unet* net, const void* x, int B, int H, int W) {
    int rc = check_shape(net, B, H, W);
    if (rc) return rc;
    net->B = B;
    net->H = H;
    net->W = W;
    net->x_in = x;
    const float s = net->cfg.leaky_slope;
    hipStream_t st = net->st;
    auto lv = [&](int level, int& h, int& w) {
        h = H >> level;
        w = W >> level;
    };
    auto pool = [&](int in, int out, int level_out) {
        int h, w;
        lv(level_out, h, w);
        const int C = TENSORS[in].ch;
        hipLaunchKernelGGL(k_pool, dim3(blocks((long)B * h * w * (C / 8))), dim3(256), 0, st, T(net, in),
                           TW(net, out), B, h, w, C);
    };
    for (int l = 0; l < SHK_L_HEAD; ++l) {
        int h, w;
        lv(LAYERS[l].level, h, w);
        const bf16* xa = l == 0 ? (const bf16*)x : T(net, IO[l].in_a);
        const bf16* xb = IO[l].in_b >= 0 ? T(net, IO[l].in_b) : nullptr;
        conv_fwd(l, xa, xb, net->wg[l], net->theta + layer_offset(l) + (int64_t)LAYERS[l].cout * 9 *
                                                                            (LAYERS[l].ca + LAYERS[l].cb),
                 TW(net, IO[l].out), B, h, w, s, st);
        if (l == SHK_L_ENC0B) pool(SHK_T_S0, SHK_T_P0, 1);
        if (l == SHK_L_ENC1B) pool(SHK_T_S1, SHK_T_P1, 2);
        if (l == SHK_L_ENC2B) pool(SHK_T_S2, SHK_T_P2, 3);
    }
    const long P = (long)B * H * W;
    hipLaunchKernelGGL(k_head, dim3(blocks(P)), dim3(256), 0, st, T(net, SHK_T_V0),
                       net->theta + layer_offset(SHK_L_HEAD), (float*)net->t[SHK_T_OUT], P);
    NNCHK(hipGetLastError());
    return SHK_OK;
}

float* output(shk_unet* net) { return (float*)net->t[SHK_T_OUT]; }
void set_stream(shk_unet* net, hipStream_t st) { net->st = st ? st : net->own; }
// back to the network's own stream, if it still runs on st
void release_stream(shk_unet* net, hipStream_t st) {
    if (net->st == st) net->st = net->own;
}
bool fits(const shk_unet* net, int H, int W) {
    return net->cfg.max_batch >= 1 && H <= net->cfg.max_h && W <= net->cfg.max_w;
}

int backward(shk_unet* net, const float* dout) {
    if (!net->x_in) return nn_fail(net, SHK_ERR_STATE, "backward before forward");
    const int B = net->B, H = net->H, W = net->W;
    const float s = net->cfg.leaky_slope;
    hipStream_t st = net->st;
    const long P0 = (long)B * H * W;
    if (dout && dout != (const float*)net->t[SHK_T_DOUT])
        NNCHK(hipMemcpyAsync(net->t[SHK_T_DOUT], dout, P0 * 4, hipMemcpyDeviceToDevice, st));
    // head
    {
        long S = P0 / 8192;
        if (S < 1) S = 1;
        if (S > 512) S = 512;
        const long nps = (P0 + S - 1) / S;
        hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)S), dim3(256), 0, st, T(net, SHK_T_V0),
                           net->theta + layer_offset(SHK_L_HEAD), (const float*)net->t[SHK_T_DOUT],
                           TW(net, SHK_T_DV0), P0, nps, net->part);
        hipLaunchKernelGGL(k_reduce_dw, dim3(1), dim3(256), 0, st, net->part, (int)S, 1, 17, 17,
                           net->grad + layer_offset(SHK_L_HEAD));
    }
    auto act_grad = [&](int g, int y) {
        int h = H >> TENSORS[y].level, w = W >> TENSORS[y].level;
        const long n8 = (long)B * h * w * TENSORS[y].ch / 8;
        hipLaunchKernelGGL(k_act_grad, dim3(blocks(n8)), dim3(256), 0, st, TW(net, g), T(net, y), n8, s);
    };
    for (int l = SHK_L_HEAD - 1; l >= 0; --l) {
        const LayerDef& L = LAYERS[l];
        const LayerIO& io = IO[l];
        const int h = H >> L.level, w = W >> L.level;
        const long P = (long)B * h * w;
        act_grad(io.d_out, io.out);
        const bf16* xa = l == 0 ? (const bf16*)net->x_in : T(net, io.in_a);
        const bf16* xb = io.in_b >= 0 ? T(net, io.in_b) : nullptr;
        float* gw = net->grad + layer_offset(l);
        conv_dw(l, xa, xb, T(net, io.d_out), net->part, net->part_cap, gw, B, h, w, st);
        bias_grad(L.cout, T(net, io.d_out), P, net->part, gw + (int64_t)L.cout * 9 * (L.ca + L.cb), st);
        if (l == 0) break;
        conv_bwd_data(l, T(net, io.d_out), net->wt[l], TW(net, io.d_in), B, h, w, st);
        if (L.up) {
            // split the concatenation gradient: deeper tensor (upsampled) and skip
            const int deep = io.in_a, skip = io.in_b;
            const int dd = deep == SHK_T_T3 ? SHK_T_DT3 : (deep == SHK_T_V2 ? SHK_T_DV2 : SHK_T_DV1);
            const int ds = skip == SHK_T_S2 ? SHK_T_DS2 : (skip == SHK_T_S1 ? SHK_T_DS1 : SHK_T_DS0);
            const int hc = h / 2, wc = w / 2;
            hipLaunchKernelGGL(k_split_up, dim3(blocks((long)B * hc * wc * (L.ca / 8))), dim3(256), 0, st,
                               T(net, io.d_in), TW(net, dd), B, hc, wc, L.ca, L.ca + L.cb);
            hipLaunchKernelGGL(k_split_skip, dim3(blocks(P * (L.cb / 8))), dim3(256), 0, st, T(net, io.d_in),
                               TW(net, ds), P, L.ca, L.cb);
        }
        // the first conv of encoder levels 1..3 feeds from a pooled skip: add
        // the pooling's backward into that skip's gradient
        if (l == SHK_L_ENC1A || l == SHK_L_ENC2A || l == SHK_L_ENC3A) {
            const int ds = l == SHK_L_ENC1A ? SHK_T_DS0 : (l == SHK_L_ENC2A ? SHK_T_DS1 : SHK_T_DS2);
            const int hf = h * 2, wf = w * 2;
            const int C = TENSORS[ds].ch;
            hipLaunchKernelGGL(k_pool_bwd_add, dim3(blocks((long)B * hf * wf * (C / 8))), dim3(256), 0, st,
                               T(net, io.d_in), TW(net, ds), B, hf, wf, C);
        }
    }
    NNCHK(hipGetLastError());
    return SHK_OK;
}

int loss(shk_unet* net, const float* tgt, int with_grad, double* out) {
    if (!net->x_in) return nn_fail(net, SHK_ERR_STATE, "loss before forward");
    const int B = net->B;
    const long HW = (long)net->H * net->W;
    long nblk = HW / 16384;
    if (nblk < 1) nblk = 1;
    if (nblk > 64) nblk = 64;
    const long nps = (HW + nblk - 1) / nblk;
    hipStream_t st = net->st;
    if (net->cfg.loss_kind == 1) {
        const bf16* x = (const bf16*)net->x_in;
        hipLaunchKernelGGL(k_eloss_sums, dim3((unsigned)nblk, B), dim3(256), 0, st,
                           (const float*)net->t[SHK_T_OUT], x, tgt, net->H, net->W, nps, net->lpart);
        hipLaunchKernelGGL(k_eloss_finish, dim3(1), dim3(256), 0, st, net->lpart, (int)nblk, B, net->lcoef);
        if (with_grad)
            hipLaunchKernelGGL(k_eloss_grad, dim3(blocks(HW * B)), dim3(256), 0, st,
                               (const float*)net->t[SHK_T_OUT], x, net->lcoef, net->H, net->W, B,
                               (float*)net->t[SHK_T_DOUT]);
    } else {
        hipLaunchKernelGGL(k_loss_sums, dim3((unsigned)nblk, B), dim3(256), 0, st,
                           (const float*)net->t[SHK_T_OUT], tgt, HW, nps, net->lpart);
        hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(256), 0, st, net->lpart, (int)nblk, B, net->lcoef);
        if (with_grad)
            hipLaunchKernelGGL(k_loss_grad, dim3(blocks(HW * B)), dim3(256), 0, st,
                               (const float*)net->t[SHK_T_OUT], tgt, net->lcoef, HW, B, (float*)net->t[SHK_T_DOUT]);
    }
    NNCHK(hipGetLastError());
    std::vector<double> c(3 * B);
    NNCHK(hipMemcpyAsync(c.data(), net->lcoef, 3 * B * sizeof(double), hipMemcpyDeviceToHost, st));
    NNCHK(hipStreamSynchronize(st));
    double s = 0;
    for (int b = 0; b < B; ++b) s += c[3 * b];
    if (out) *out = s / B;
    return SHK_OK;
}

const char* unet_error(const shk_unet* net) { return net->err.c_str(); }

}  // namespace shknn

// ------------------------------------------------------------------ C ABI

extern "C" {

int shk_dev_alloc(size_t bytes, void** out) {
    if (!out) return SHK_ERR_ARG;
    *out = nullptr;
    if (bytes == 0) return SHK_ERR_ARG;
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return SHK_ERR_HIP;
    }
    return SHK_OK;
}

int shk_dev_free(void* ptr) { return hipFree(ptr) == hipSuccess ? SHK_OK : SHK_ERR_HIP; }

int shk_dev_copy(void* dst, const void* src, size_t bytes, int32_t kind) {
    if ((!dst || !src) && bytes) return SHK_ERR_ARG;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                      : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (kind < 0 || kind > 2) return SHK_ERR_ARG;
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? SHK_OK : SHK_ERR_HIP;
}

int64_t shk_unet_num_params(void) { return NUM_PARAMS; }

int shk_unet_layer_info(int32_t layer, int64_t* offset, int32_t* cin, int32_t* cout) {
    if (layer < 0 || layer >= SHK_L_COUNT) return SHK_ERR_ARG;
    if (offset) *offset = layer_offset(layer);
    if (cin) *cin = LAYERS[layer].ca + LAYERS[layer].cb;
    if (cout) *cout = LAYERS[layer].cout;
    return SHK_OK;
}

const char* shk_unet_last_error(const shk_unet* net) { return net ? net->err.c_str() : "null network"; }

int shk_unet_create(const shk_unet_cfg* cfg, shk_unet** out) {
    if (!cfg || !out) return SHK_ERR_ARG;
    *out = nullptr;
    if (cfg->max_batch < 1 || cfg->max_batch > 256 || cfg->max_h < 32 || cfg->max_w < 64 || cfg->max_h % 32 ||
        cfg->max_w % 64 || cfg->loss_kind < 0 || cfg->loss_kind > 1 ||
        !(cfg->leaky_slope >= 0.f && cfg->leaky_slope < 1.f))
        return SHK_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SHK_ERR_NODEVICE;
    shk_unet* net = new shk_unet();
    net->cfg = *cfg;
    auto bail = [&](int rc) {
        shk_unet_destroy(net);
        return rc;
    };
    if (hipStreamCreateWithFlags(&net->own, hipStreamNonBlocking) != hipSuccess) return bail(SHK_ERR_HIP);
    net->st = net->own;
    const size_t np = (size_t)NUM_PARAMS;
    float** f4[] = {&net->theta, &net->grad, &net->m, &net->v};
    for (auto pp : f4) {
        if (hipMalloc((void**)pp, np * 4 + 256) != hipSuccess) return bail(SHK_ERR_HIP);
        if (hipMemset(*pp, 0, np * 4 + 256) != hipSuccess) return bail(SHK_ERR_HIP);
    }
    for (int l = 0; l < SHK_L_HEAD; ++l) {
        const int cin = LAYERS[l].ca + LAYERS[l].cb, cout = LAYERS[l].cout;
        if (hipMalloc((void**)&net->wg[l], (size_t)cout * kpad(cin) * 2 + 256) != hipSuccess) return bail(SHK_ERR_HIP);
        if (hipMalloc((void**)&net->wt[l], (size_t)cin * kpad(cout) * 2 + 256) != hipSuccess) return bail(SHK_ERR_HIP);
    }
    const size_t P0 = (size_t)cfg->max_batch * cfg->max_h * cfg->max_w;
    for (int i = 0; i < SHK_T_COUNT; ++i) {
        const size_t px = P0 >> (2 * TENSORS[i].level);
        net->t_elems[i] = px * TENSORS[i].ch;
        if (i == SHK_T_X) continue;  // the input stays in the caller's buffer
        const size_t bytes = net->t_elems[i] * (TENSORS[i].f32 ? 4 : 2) + 256;
        if (hipMalloc(&net->t[i], bytes) != hipSuccess) return bail(SHK_ERR_HIP);
        if (hipMemset(net->t[i], 0, bytes) != hipSuccess) return bail(SHK_ERR_HIP);
    }
    // weight-gradient partials: 2048 workgroup tiles of at most 32 x 32, or
    // one per 4096 pixels for the bias / head sums
    net->part_cap = std::max((size_t)2048 * 32 * 32, (size_t)512 * 128) + 17 * 512;
    if (hipMalloc((void**)&net->part, net->part_cap * 4) != hipSuccess) return bail(SHK_ERR_HIP);
    net->lpart_cap = (size_t)cfg->max_batch * 64 * 3;
    if (hipMalloc((void**)&net->lpart, net->lpart_cap * 8) != hipSuccess) return bail(SHK_ERR_HIP);
    if (hipMalloc((void**)&net->lcoef, (size_t)cfg->max_batch * 3 * 8 + 256) != hipSuccess) return bail(SHK_ERR_HIP);
    if (pack_weights(net)) return bail(SHK_ERR_HIP);
    if (hipStreamSynchronize(net->st) != hipSuccess) return bail(SHK_ERR_HIP);
    *out = net;
    return SHK_OK;
}

int shk_unet_destroy(shk_unet* net) {
    if (!net) return SHK_OK;
    if (net->own) hipStreamSynchronize(net->own);
    hipFree(net->theta);
    hipFree(net->grad);
    hipFree(net->m);
    hipFree(net->v);
    for (int l = 0; l < SHK_L_COUNT; ++l) {
        hipFree(net->wg[l]);
        hipFree(net->wt[l]);
    }
    for (int i = 0; i < SHK_T_COUNT; ++i) hipFree(net->t[i]);
    hipFree(net->part);
    hipFree(net->lpart);
    hipFree(net->lcoef);
    if (net->own) hipStreamDestroy(net->own);
    delete net;
    return SHK_OK;
}

int shk_unet_set_params(shk_unet* net, const float* host, int64_t n) {
    if (!net || !host || n != NUM_PARAMS) return nn_fail(net, SHK_ERR_ARG, "parameter count mismatch");
    NNCHK(hipMemcpyAsync(net->theta, host, n * 4, hipMemcpyHostToDevice, net->st));
    NNCHK(hipMemsetAsync(net->m, 0, n * 4, net->st));
    NNCHK(hipMemsetAsync(net->v, 0, n * 4, net->st));
    net->adam_t = 0;
    int rc = pack_weights(net);
    if (rc) return rc;
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_get_params(shk_unet* net, float* host, int64_t n) {
    if (!net || !host || n != NUM_PARAMS) return nn_fail(net, SHK_ERR_ARG, "parameter count mismatch");
    NNCHK(hipMemcpyAsync(host, net->theta, n * 4, hipMemcpyDeviceToHost, net->st));
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_get_grads(shk_unet* net, float* host, int64_t n) {
    if (!net || !host || n != NUM_PARAMS) return nn_fail(net, SHK_ERR_ARG, "parameter count mismatch");
    NNCHK(hipMemcpyAsync(host, net->grad, n * 4, hipMemcpyDeviceToHost, net->st));
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_init_params(shk_unet* net, uint64_t seed) {
    if (!net) return SHK_ERR_ARG;
    // He-normal weights (std sqrt(2 / fan_in)), zero biases; splitmix64 +
    // Box-Muller on the host so the draw is the same everywhere
    std::vector<float> th((size_t)NUM_PARAMS, 0.f);
    uint64_t state = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    auto next = [&]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < SHK_L_COUNT; ++l) {
        const int cin = LAYERS[l].ca + LAYERS[l].cb, cout = LAYERS[l].cout;
        const int fan_in = l == SHK_L_HEAD ? 16 : 9 * cin;
        const int64_t nw = l == SHK_L_HEAD ? 16 : (int64_t)cout * 9 * cin;
        const double sd = std::sqrt((l == SHK_L_HEAD ? 1.0 : 2.0) / fan_in);
        for (int64_t i = 0; i < nw; ++i) {
            const double u1 = next(), u2 = next();
            th[(size_t)(layer_offset(l) + i)] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    return shk_unet_set_params(net, th.data(), NUM_PARAMS);
}

int shk_unet_forward(shk_unet* net, const void* x_dev, int32_t B, int32_t H, int32_t W, float* y_dev) {
    if (!net || !x_dev) return nn_fail(net, SHK_ERR_ARG, "null argument");
    int rc = forward(net, x_dev, B, H, W);
    if (rc) return rc;
    if (y_dev) NNCHK(hipMemcpyAsync(y_dev, net->t[SHK_T_OUT], (size_t)B * H * W * 4, hipMemcpyDeviceToDevice, net->st));
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_loss(shk_unet* net, const float* t_dev, int32_t with_grad, double* out) {
    if (!net || !t_dev) return nn_fail(net, SHK_ERR_ARG, "null argument");
    return loss(net, t_dev, with_grad, out);
}

int shk_unet_backward(shk_unet* net, const float* dout_dev) {
    if (!net) return SHK_ERR_ARG;
    int rc = backward(net, dout_dev);
    if (rc) return rc;
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_adam(shk_unet* net, float lr, float beta1, float beta2, float eps) {
    if (!net) return SHK_ERR_ARG;
    net->adam_t += 1;
    const float bc1 = 1.f - std::pow(beta1, (float)net->adam_t);
    const float bc2 = 1.f - std::pow(beta2, (float)net->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(NUM_PARAMS)), dim3(256), 0, net->st, net->theta, net->grad, net->m,
                       net->v, (long)NUM_PARAMS, lr, beta1, beta2, eps, bc1, bc2);
    int rc = pack_weights(net);
    if (rc) return rc;
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

int shk_unet_train_step(shk_unet* net, const void* x_dev, const float* t_dev, int32_t B, int32_t H, int32_t W,
                        float lr, double* out) {
    if (!net || !x_dev || !t_dev) return nn_fail(net, SHK_ERR_ARG, "null argument");
    int rc = forward(net, x_dev, B, H, W);
    if (!rc) rc = loss(net, t_dev, 1, out);
    if (!rc) rc = backward(net, nullptr);
    if (!rc) rc = shk_unet_adam(net, lr, 0.9f, 0.999f, 1e-8f);
    return rc;
}

int shk_unet_download_tensor(shk_unet* net, int32_t id, float* host, int64_t count) {
    if (!net || !host || id < 0 || id >= SHK_T_COUNT) return nn_fail(net, SHK_ERR_ARG, "bad tensor id");
    if (!net->x_in) return nn_fail(net, SHK_ERR_STATE, "no pass has run");
    const int lvl = TENSORS[id].level;
    const long n = (long)net->B * (net->H >> lvl) * (net->W >> lvl) * TENSORS[id].ch;
    if (count != n) return nn_fail(net, SHK_ERR_ARG, "count does not match the tensor of the last pass");
    if (TENSORS[id].f32) {
        NNCHK(hipMemcpyAsync(host, net->t[id], n * 4, hipMemcpyDeviceToHost, net->st));
    } else {
        float* tmp = nullptr;
        NNCHK(hipMalloc((void**)&tmp, n * 4));
        const void* src = id == SHK_T_X ? net->x_in : net->t[id];
        hipLaunchKernelGGL(k_to_float, dim3(blocks(n)), dim3(256), 0, net->st, (const bf16*)src, tmp, n);
        hipError_t e = hipMemcpyAsync(host, tmp, n * 4, hipMemcpyDeviceToHost, net->st);
        hipStreamSynchronize(net->st);
        hipFree(tmp);
        if (e != hipSuccess) return nn_fail(net, SHK_ERR_HIP, hipGetErrorString(e));
    }
    NNCHK(hipStreamSynchronize(net->st));
    return SHK_OK;
}

}  // extern "C"
