// The synthetic glacier of the benchmark configurations, generated on the
// device (shk_generate_glacier).
//
// Every value is built from IEEE basic operations only (integer hashing,
// exact sums of 48-bit fractions, separately rounded products and sums in a
// fixed order); the few transcendental values -- the smoothing weights, the
// trough profile, the surface profile, the standard deviation -- are
// computed on the host, so a CPU restatement with the same loops reproduces
// the fields bit for bit.
//
//   stream   o_i = splitmix64(seed + i * phi), i = 0, 1, ...: the outputs of
//            Vigna's SplitMix64 generator started from state `seed`
//            (phi = 0x9E3779B97F4A7C15);
//   noise    w_c = sum_{k<12} (o_{12c+k} >> 16) 2^-48 - 6 for cell
//            c = ix + iy nx: Irwin-Hall(12), mean 0, variance 1, exact;
//   smooth   periodic convolution along x, then along y, with the weights
//            exp(-(k/s)^2 / 2) / sum, k = -R..R, s = corr_len / (sqrt(2) d),
//            R = ceil(4 s) (no smoothing for corr_len 0): the field's
//            covariance is exp(-r^2 / (2 corr_len^2)) up to the truncation;
//   scale    zb = (f - mean) * bed_amp / sd, mean and sd over the grid, the
//            sums taken in 16384 interleaved chunks (chunk j holds the cells
//            j, j + 16384, ...), each summed in order, then 128 groups of 128
//            consecutive chunk sums, each in order, then the group sums in
//            order;
//   trough   zb -= depth exp(-((iy dy - y0) / (width / 2))^2 / 2),
//            y0 = (1/4 + u / 2) ly, u = (o_{12n} >> 11) 2^-53;
//   surface  zs = 6 (sqrt(ix dx + 5000) - sqrt(5000)) + 1;
//   moulins  from o_{12n+1} on: a cell of the interior (1..nx-2) x
//            (1..ny-2) drawn as floor(u M), u = (o >> 11) 2^-53, drawn again
//            while already taken, then Q = q_min + (q_max - q_min) u from
//            the next output; moulin = Q / (dx dy) there, 0 elsewhere.

#include <math.h>
#include <string.h>

#include <string>
#include <unordered_set>
#include <vector>

#include "shk_internal.hpp"

namespace shkgen {

using shk::Fields;
using shk::GROUNDED;
using shk::OCEAN;

namespace {

constexpr uint64_t PHI = 0x9E3779B97F4A7C15ull;
constexpr int NCHUNK = 16384;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += PHI;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gen_noise(size_t n, uint64_t seed, double* __restrict__ w) {
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const uint64_t base = seed + (uint64_t)(12 * c) * PHI;
        double s = 0.0;
        for (int k = 0; k < 12; ++k) s = s + (double)(splitmix64(base + (uint64_t)k * PHI) >> 16) * 0x1p-48;
        w[c] = s - 6.0;
    }
}

__global__ __launch_bounds__(256) void k_gen_conv_x(int nx, int ny, int R, const double* __restrict__ wts,
                                                    const double* __restrict__ in, double* __restrict__ out) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    const double* row = in + (size_t)iy * nx;
    int j = (ix - R) % nx;
    if (j < 0) j += nx;
    double acc = 0.0;
    for (int k = 0; k <= 2 * R; ++k) {
        acc = acc + wts[k] * row[j];
        if (++j == nx) j = 0;
    }
    out[(size_t)iy * nx + ix] = acc;
}

__global__ __launch_bounds__(256) void k_gen_conv_y(int nx, int ny, int R, const double* __restrict__ wts,
                                                    const double* __restrict__ in, double* __restrict__ out) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y;
    if (ix >= nx || iy >= ny) return;
    int j = (iy - R) % ny;
    if (j < 0) j += ny;
    double acc = 0.0;
    for (int k = 0; k <= 2 * R; ++k) {
        acc = acc + wts[k] * in[(size_t)j * nx + ix];
        if (++j == ny) j = 0;
    }
    out[(size_t)iy * nx + ix] = acc;
}

// chunk j: f[j] + f[j + NCHUNK] + ... in order; with square, of (f - mean)^2
__global__ __launch_bounds__(256) void k_gen_chunks(size_t n, const double* __restrict__ f, double mean, int square,
                                                    double* __restrict__ part) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= NCHUNK) return;
    double acc = 0.0;
    for (size_t i = c; i < n; i += NCHUNK) {
        double v = f[i];
        if (square) {
            v = v - mean;
            v = v * v;
        }
        acc = acc + v;
    }
    part[c] = acc;
}

// the chunk sums: lane t sums chunks 128 t .. 128 t + 127 in order, lane 0
// the 128 group sums in order
__global__ __launch_bounds__(128) void k_gen_total(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double grp[128];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int c = 0; c < 128; ++c) acc = acc + part[128 * t + c];
    grp[t] = acc;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int k = 0; k < 128; ++k) s = s + grp[k];
        out[0] = s;
    }
}

struct Uniform {
    double A, b0, G, ieb, ub;
};

__global__ __launch_bounds__(256) void k_gen_finish(int nx, int ny, const double* __restrict__ f, double mean,
                                                    double scale, int do_scale, const double* __restrict__ trough,
                                                    double depth, const double* __restrict__ zs_row, Uniform u,
                                                    Fields F) {
    const int ix = blockIdx.x * 64 + threadIdx.x, iy = blockIdx.y * 4 + threadIdx.y;
    if (ix < nx && iy < ny) {
        const size_t c = (size_t)ix + (size_t)iy * nx;
        const double d = f[c] - mean;
        double z = do_scale ? d * scale : d;
        if (trough) z = z - depth * trough[iy];
        F.f[SHK_ZB][c] = z;
        F.f[SHK_ZS][c] = zs_row[ix];
        F.f[SHK_MASK][c] = ix == 0 ? OCEAN : GROUNDED;
        F.f[SHK_A_VISC][c] = u.A;
        F.f[SHK_B][c] = u.b0;
        F.f[SHK_G][c] = u.G;
        F.f[SHK_IEB][c] = u.ieb;
    }
    if (ix <= nx && iy < ny) {
        const size_t k = (size_t)ix + (size_t)iy * (nx + 1);
        F.f[SHK_UB_X][k] = u.ub;
        F.f[SHK_TAUB_X][k] = 0.0;
    }
    if (ix < nx && iy <= ny) {
        const size_t k = (size_t)ix + (size_t)iy * nx;
        F.f[SHK_UB_Y][k] = 0.0;
        F.f[SHK_TAUB_Y][k] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_gen_zero(size_t n, double* __restrict__ a) {
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) a[c] = 0.0;
}
// one thread per moulin; the cells are distinct, so no two threads write one cell
__global__ __launch_bounds__(256) void k_gen_moulins(int m, const int64_t* __restrict__ cell,
                                                     const double* __restrict__ val, double* __restrict__ moulin) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < m) moulin[cell[k]] = val[k];
}

double unit53(uint64_t o) { return (double)(o >> 11) * 0x1p-53; }

// exp(-(k/s)^2 / 2) normalised by its sum over k = -R..R (k in order)
std::vector<double> weights(double s, int* R) {
    if (!(s > 0.0)) {
        *R = 0;
        return {1.0};
    }
    *R = (int)ceil(4.0 * s);
    std::vector<double> w(2 * *R + 1);
    double sum = 0.0;
    for (int k = -*R; k <= *R; ++k) {
        const double t = (double)k / s;
        w[k + *R] = exp(-0.5 * t * t);
        sum = sum + w[k + *R];
    }
    for (double& v : w) v = v / sum;
    return w;
}

int blocks(size_t n) {
    const size_t b = (n + 255) / 256;
    return (int)(b < 2048 ? (b ? b : 1) : 2048);
}

#define GENCHK(call)                                                  \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) {                                       \
            err = std::string(#call) + ": " + hipGetErrorString(e_); \
            goto done;                                                \
        }                                                             \
    } while (0)

}  // namespace

int generate(const shk_grid& g, const shk_glacier_spec& sp, Fields F, double* w, double* tmp, double* moulin,
             hipStream_t st, std::string& err) {
    const int nx = g.nx, ny = g.ny;
    const size_t n = (size_t)nx * ny;
    int Rx, Ry;
    const std::vector<double> wx = weights(sp.corr_len / (sqrt(2.0) * g.dx), &Rx);
    const std::vector<double> wy = weights(sp.corr_len / (sqrt(2.0) * g.dy), &Ry);
    const bool has_trough = sp.trough_width > 0.0 && sp.trough_depth != 0.0;
    std::vector<double> trough, zs(nx);
    for (int ix = 0; ix < nx; ++ix) zs[ix] = 6.0 * (sqrt((double)ix * g.dx + 5000.0) - sqrt(5000.0)) + 1.0;
    uint64_t next = 12 * (uint64_t)n;  // stream index after the noise
    if (has_trough) {
        const double y0 = (0.25 + 0.5 * unit53(splitmix64(sp.seed + next * PHI))) * g.ly;
        const double sig = sp.trough_width / 2.0;
        trough.resize(ny);
        for (int iy = 0; iy < ny; ++iy) {
            const double t = ((double)iy * g.dy - y0) / sig;
            trough[iy] = exp(-0.5 * t * t);
        }
    }
    ++next;
    std::vector<int64_t> cells;
    std::vector<double> qs;
    if (sp.n_moulins > 0) {
        const int64_t mx = nx - 2, M = (int64_t)(nx - 2) * (ny - 2);
        std::unordered_set<int64_t> taken;
        taken.reserve((size_t)sp.n_moulins * 2);
        for (int k = 0; k < sp.n_moulins; ++k) {
            int64_t c;
            for (;;) {
                const int64_t j = (int64_t)(unit53(splitmix64(sp.seed + next++ * PHI)) * (double)M);
                c = (1 + j % mx) + (1 + j / mx) * (int64_t)nx;
                if (taken.insert(c).second) break;
            }
            const double u = unit53(splitmix64(sp.seed + next++ * PHI));
            cells.push_back(c);
            qs.push_back((sp.q_min + (sp.q_max - sp.q_min) * u) / (g.dx * g.dy));
        }
    }
    // one device block for the small arrays: weights, profiles, chunk sums, moulins
    const size_t nsmall = wx.size() + wy.size() + zs.size() + trough.size() + NCHUNK + 1 + qs.size();
    double* d = nullptr;
    int64_t* dcell = nullptr;
    double* host = nullptr;
    int rc = SHK_ERR_HIP;
    {
        GENCHK(hipMalloc((void**)&d, nsmall * sizeof(double)));
        if (!cells.empty()) GENCHK(hipMalloc((void**)&dcell, cells.size() * sizeof(int64_t)));
        GENCHK(hipHostMalloc((void**)&host, 2 * sizeof(double), hipHostMallocDefault));
        double* dwx = d;
        double* dwy = dwx + wx.size();
        double* dzs = dwy + wy.size();
        double* dtr = dzs + zs.size();
        double* part = dtr + trough.size();
        double* total = part + NCHUNK;
        double* dq = total + 1;
        GENCHK(hipMemcpyAsync(dwx, wx.data(), wx.size() * 8, hipMemcpyHostToDevice, st));
        GENCHK(hipMemcpyAsync(dwy, wy.data(), wy.size() * 8, hipMemcpyHostToDevice, st));
        GENCHK(hipMemcpyAsync(dzs, zs.data(), zs.size() * 8, hipMemcpyHostToDevice, st));
        if (has_trough) GENCHK(hipMemcpyAsync(dtr, trough.data(), trough.size() * 8, hipMemcpyHostToDevice, st));
        if (!cells.empty()) {
            GENCHK(hipMemcpyAsync(dcell, cells.data(), cells.size() * 8, hipMemcpyHostToDevice, st));
            GENCHK(hipMemcpyAsync(dq, qs.data(), qs.size() * 8, hipMemcpyHostToDevice, st));
        }
        const dim3 cgrid((nx + 63) / 64, (ny + 3) / 4), blk(64, 4);
        hipLaunchKernelGGL(k_gen_noise, dim3(blocks(n)), dim3(256), 0, st, n, sp.seed, w);
        hipLaunchKernelGGL(k_gen_conv_x, cgrid, blk, 0, st, nx, ny, Rx, dwx, w, tmp);
        hipLaunchKernelGGL(k_gen_conv_y, cgrid, blk, 0, st, nx, ny, Ry, dwy, tmp, w);
        hipLaunchKernelGGL(k_gen_chunks, dim3(NCHUNK / 256), dim3(256), 0, st, n, w, 0.0, 0, part);
        hipLaunchKernelGGL(k_gen_total, dim3(1), dim3(128), 0, st, part, total);
        GENCHK(hipGetLastError());
        GENCHK(hipMemcpyAsync(host, total, 8, hipMemcpyDeviceToHost, st));
        GENCHK(hipStreamSynchronize(st));
        const double mean = host[0] / (double)n;
        hipLaunchKernelGGL(k_gen_chunks, dim3(NCHUNK / 256), dim3(256), 0, st, n, w, mean, 1, part);
        hipLaunchKernelGGL(k_gen_total, dim3(1), dim3(128), 0, st, part, total);
        GENCHK(hipGetLastError());
        GENCHK(hipMemcpyAsync(host + 1, total, 8, hipMemcpyDeviceToHost, st));
        GENCHK(hipStreamSynchronize(st));
        const double sd = sqrt(host[1] / (double)n);
        const Uniform u{sp.A_visc, sp.b0, sp.G, sp.ieb, sp.ub};
        hipLaunchKernelGGL(k_gen_finish, dim3((nx + 64) / 64, (ny + 4) / 4), blk, 0, st, nx, ny, w, mean,
                           sd > 0.0 ? sp.bed_amp / sd : 1.0, sd > 0.0 ? 1 : 0, has_trough ? dtr : nullptr,
                           sp.trough_depth, dzs, u, F);
        if (moulin && !cells.empty()) {
            hipLaunchKernelGGL(k_gen_zero, dim3(blocks(n)), dim3(256), 0, st, n, moulin);
            hipLaunchKernelGGL(k_gen_moulins, dim3((unsigned)((cells.size() + 255) / 256)), dim3(256), 0, st,
                               (int)cells.size(), dcell, dq, moulin);
        }
        GENCHK(hipGetLastError());
        GENCHK(hipStreamSynchronize(st));
        rc = SHK_OK;
    }
done:
    if (d) hipFree(d);
    if (dcell) hipFree(dcell);
    if (host) hipHostFree(host);
    return rc;
}

}  // namespace shkgen
