// Checkpoint / restart and field statistics of the device State.
//
// Reference: save_checkpoint / load_checkpoint! (src/checkpoint.jl:26-64)
// write every State field plus t and total_time to a file that replaces the
// previous one only once complete (written to path.tmp, then renamed), and
// restore them in place.  The observer's CSV writer records min / max / mean
// of each tracked field (src/observer.jl:324-339).
//
// Here a checkpoint is taken in two halves so the solver is held up only for
// a device-to-device copy: `begin` copies the fields into a device staging
// area on the compute stream (the State may change as soon as that copy is
// queued), computes one checksum per field there, and queues the copy to
// pinned host memory on a second stream; the file is then written from the
// pinned copy, on a host thread when asked, while later steps run.
//
// File layout (little-endian, every block 512-byte aligned):
//   Header (64 bytes): magic "SHKCKPT1", version, nfields, nx, ny, t,
//     total_time, lx, ly, data_offset
//   nfields Entry records (64 bytes): field name (the reference State's),
//     field id, rows, cols, byte offset, byte count, checksum
//   the fields, column-major (ix fastest), FP64.
// The checksum of a field of n FP64 values with bit patterns w_i is
//   sum_i splitmix64(w_i + i * 0x9E3779B97F4A7C15) mod 2^64,
// splitmix64 being Vigna's SplitMix64 finaliser; the sum is
// order-independent, so it is reduced on the device, and the position term
// makes swapped or shifted blocks fail it.

#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "shakti_hip.h"
#include "shk_io.hpp"

namespace shkio {

namespace {

constexpr char kMagic[8] = {'S', 'H', 'K', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kAlignWords = 64;  // 512 bytes
constexpr int kBlocks = 512;
constexpr int kThreads = 256;
constexpr int kMaxFields = SHK_NUM_FIELDS;

struct Header {
    char magic[8];
    uint32_t version, nfields;
    int32_t nx, ny;
    int64_t t;
    double total_time, lx, ly;
    uint64_t data_offset;
};
static_assert(sizeof(Header) == 64, "checkpoint header is 64 bytes");

struct Entry {
    char name[16];
    int32_t field_id, rows, cols, reserved;
    uint64_t offset, bytes, checksum, reserved2;
};
static_assert(sizeof(Entry) == 64, "checkpoint entry is 64 bytes");

size_t align_up(size_t w) { return (w + kAlignWords - 1) / kAlignWords * kAlignWords; }

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kThreads) void k_mix_part(const uint64_t* __restrict__ w, size_t n,
                                                       unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sh[kThreads];
    unsigned long long acc = 0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        acc += splitmix64(w[i] + (uint64_t)i * 0x9E3779B97F4A7C15ull);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kThreads) void k_mix_final(const unsigned long long* __restrict__ part, int nb,
                                                        unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sh[kThreads];
    unsigned long long acc = 0;
    for (int i = threadIdx.x; i < nb; i += kThreads) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// per block: min, max, sum and a NaN flag of its share of x
__global__ __launch_bounds__(kThreads) void k_stats_part(const double* __restrict__ x, size_t n,
                                                         double* __restrict__ part) {
    __shared__ double smin[kThreads], smax[kThreads], ssum[kThreads], snan[kThreads];
    double mn = INFINITY, mx = -INFINITY, sm = 0.0, nan = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const double v = x[i];
        if (v != v) {
            nan = 1.0;
        } else {
            mn = fmin(mn, v);
            mx = fmax(mx, v);
        }
        sm += v;
    }
    const int t = threadIdx.x;
    smin[t] = mn;
    smax[t] = mx;
    ssum[t] = sm;
    snan[t] = nan;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            smin[t] = fmin(smin[t], smin[t + s]);
            smax[t] = fmax(smax[t], smax[t + s]);
            ssum[t] += ssum[t + s];
            snan[t] = fmax(snan[t], snan[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[4 * blockIdx.x + 0] = smin[0];
        part[4 * blockIdx.x + 1] = smax[0];
        part[4 * blockIdx.x + 2] = ssum[0];
        part[4 * blockIdx.x + 3] = snan[0];
    }
}

__global__ __launch_bounds__(kThreads) void k_stats_final(const double* __restrict__ part, int nb, double count,
                                                          double* __restrict__ out) {
    __shared__ double smin[kThreads], smax[kThreads], ssum[kThreads], snan[kThreads];
    double mn = INFINITY, mx = -INFINITY, sm = 0.0, nan = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
        mn = fmin(mn, part[4 * b + 0]);
        mx = fmax(mx, part[4 * b + 1]);
        sm += part[4 * b + 2];
        nan = fmax(nan, part[4 * b + 3]);
    }
    const int t = threadIdx.x;
    smin[t] = mn;
    smax[t] = mx;
    ssum[t] = sm;
    snan[t] = nan;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            smin[t] = fmin(smin[t], smin[t + s]);
            smax[t] = fmax(smax[t], smax[t + s]);
            ssum[t] += ssum[t + s];
            snan[t] = fmax(snan[t], snan[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const bool any_nan = snan[0] != 0.0;
        out[0] = any_nan ? NAN : smin[0];
        out[1] = any_nan ? NAN : smax[0];
        out[2] = ssum[0] / count;
    }
}

int blocks_for(size_t n) {
    const size_t b = (n + kThreads - 1) / kThreads;
    return (int)(b < (size_t)kBlocks ? (b ? b : 1) : kBlocks);
}

#define IOCHK(call)                                                                 \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return SHK_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)

}  // namespace

struct Snapshot {
    hipStream_t cs = nullptr;  // copy stream, default priority
    hipEvent_t ev_staged = nullptr, ev_copied = nullptr;
    double* dev = nullptr;  // device staging
    size_t dev_cap = 0;     // in words
    double* host = nullptr; // pinned host staging
    size_t host_cap = 0;
    unsigned long long* part = nullptr;     // kBlocks checksum partials
    unsigned long long* dev_ck = nullptr;   // one checksum per field
    unsigned long long* host_ck = nullptr;  // pinned
    double* stat_part = nullptr;            // 4 * kBlocks
    double* stat_dev = nullptr;             // 3
    double* stat_host = nullptr;            // pinned 3
    std::vector<int> ids;
    std::vector<size_t> off, cnt;
    std::vector<FieldDesc> desc;
    bool have = false;    // the staging holds a snapshot (ids/off/cnt describe it)
    bool copied = false;  // ev_copied has been recorded at least once
    std::thread writer;
    bool writing = false;
    int writer_rc = SHK_OK;
    std::string writer_err;
};

Snapshot* create(std::string& err) {
    Snapshot* s = new Snapshot();
    auto bad = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        destroy(s);
        return nullptr;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking)) != hipSuccess) return bad(e, "hipStreamCreate");
    if ((e = hipEventCreateWithFlags(&s->ev_staged, hipEventDisableTiming)) != hipSuccess) return bad(e, "hipEventCreate");
    if ((e = hipEventCreateWithFlags(&s->ev_copied, hipEventDisableTiming)) != hipSuccess) return bad(e, "hipEventCreate");
    if ((e = hipMalloc((void**)&s->part, kBlocks * sizeof(unsigned long long))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&s->dev_ck, kMaxFields * sizeof(unsigned long long))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipHostMalloc((void**)&s->host_ck, kMaxFields * sizeof(unsigned long long), hipHostMallocDefault)) !=
        hipSuccess)
        return bad(e, "hipHostMalloc");
    if ((e = hipMalloc((void**)&s->stat_part, 4 * kBlocks * sizeof(double))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipMalloc((void**)&s->stat_dev, 4 * sizeof(double))) != hipSuccess) return bad(e, "hipMalloc");
    if ((e = hipHostMalloc((void**)&s->stat_host, 4 * sizeof(double), hipHostMallocDefault)) != hipSuccess)
        return bad(e, "hipHostMalloc");
    return s;
}

void destroy(Snapshot* s) {
    if (!s) return;
    if (s->writer.joinable()) s->writer.join();
    if (s->cs) hipStreamSynchronize(s->cs);
    void* dptrs[] = {s->dev, s->part, s->dev_ck, s->stat_part, s->stat_dev};
    for (void* p : dptrs)
        if (p) hipFree(p);
    void* hptrs[] = {s->host, s->host_ck, s->stat_host};
    for (void* p : hptrs)
        if (p) hipHostFree(p);
    if (s->ev_staged) hipEventDestroy(s->ev_staged);
    if (s->ev_copied) hipEventDestroy(s->ev_copied);
    if (s->cs) hipStreamDestroy(s->cs);
    delete s;
}

namespace {

void join_writer(Snapshot* s) {
    if (s->writer.joinable()) s->writer.join();
    s->writing = false;
}

// Room for `words` in both staging areas; waits for every copy that uses them.
int reserve(Snapshot* s, hipStream_t st, size_t words, std::string& err) {
    if (words <= s->dev_cap && words <= s->host_cap) return SHK_OK;
    IOCHK(hipStreamSynchronize(st));
    IOCHK(hipStreamSynchronize(s->cs));
    if (words > s->dev_cap) {
        if (s->dev) IOCHK(hipFree(s->dev));
        s->dev = nullptr;
        s->dev_cap = 0;
        IOCHK(hipMalloc((void**)&s->dev, words * sizeof(double)));
        s->dev_cap = words;
    }
    if (words > s->host_cap) {
        if (s->host) IOCHK(hipHostFree(s->host));
        s->host = nullptr;
        s->host_cap = 0;
        IOCHK(hipHostMalloc((void**)&s->host, words * sizeof(double), hipHostMallocDefault));
        s->host_cap = words;
    }
    s->have = false;
    return SHK_OK;
}

// checksum of dev[0..n) into s->dev_ck[j], on st
void checksum(Snapshot* s, hipStream_t st, const double* dev, size_t n, int j) {
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k_mix_part, dim3(nb), dim3(kThreads), 0, st, (const uint64_t*)dev, n, s->part);
    hipLaunchKernelGGL(k_mix_final, dim3(1), dim3(kThreads), 0, st, s->part, nb, s->dev_ck + j);
}

}  // namespace

int begin(Snapshot* s, hipStream_t st, int nfields, const int* ids, const FieldDesc* all, std::string& err) {
    if (nfields < 1 || nfields > kMaxFields) {
        err = "a snapshot holds 1 to " + std::to_string(kMaxFields) + " fields";
        return SHK_ERR_ARG;
    }
    for (int j = 0; j < nfields; ++j) {
        if (ids[j] < 0 || ids[j] >= kMaxFields) {
            err = "unknown field id " + std::to_string(ids[j]);
            return SHK_ERR_ARG;
        }
        for (int k = 0; k < j; ++k)
            if (ids[k] == ids[j]) {
                err = "field id " + std::to_string(ids[j]) + " listed twice";
                return SHK_ERR_ARG;
            }
    }
    join_writer(s);  // its outcome stays in writer_rc for wait()
    std::vector<size_t> off(nfields), cnt(nfields);
    size_t total = 0;
    for (int j = 0; j < nfields; ++j) {
        const FieldDesc& d = all[ids[j]];
        off[j] = total;
        cnt[j] = (size_t)d.rows * (size_t)d.cols;
        total += align_up(cnt[j]);
    }
    if (int rc = reserve(s, st, total, err)) return rc;
    // the previous snapshot's copy to the host still reads the staging area
    if (s->copied) IOCHK(hipStreamWaitEvent(st, s->ev_copied, 0));
    for (int j = 0; j < nfields; ++j) {
        IOCHK(hipMemcpyAsync(s->dev + off[j], all[ids[j]].dev, cnt[j] * sizeof(double), hipMemcpyDeviceToDevice, st));
        checksum(s, st, s->dev + off[j], cnt[j], j);
    }
    IOCHK(hipGetLastError());
    IOCHK(hipEventRecord(s->ev_staged, st));
    IOCHK(hipStreamWaitEvent(s->cs, s->ev_staged, 0));
    IOCHK(hipMemcpyAsync(s->host, s->dev, total * sizeof(double), hipMemcpyDeviceToHost, s->cs));
    IOCHK(hipMemcpyAsync(s->host_ck, s->dev_ck, nfields * sizeof(unsigned long long), hipMemcpyDeviceToHost, s->cs));
    IOCHK(hipEventRecord(s->ev_copied, s->cs));
    s->copied = true;
    s->ids.assign(ids, ids + nfields);
    s->off = off;
    s->cnt = cnt;
    s->desc.assign(all, all + kMaxFields);
    s->have = true;
    return SHK_OK;
}

int fetch(Snapshot* s, int id, double* host, size_t count, std::string& err) {
    if (!s->have) {
        err = "no snapshot has been taken";
        return SHK_ERR_STATE;
    }
    for (size_t j = 0; j < s->ids.size(); ++j) {
        if (s->ids[j] != id) continue;
        if (count != s->cnt[j]) {
            err = "field " + std::to_string(id) + " holds " + std::to_string(s->cnt[j]) + " values, not " +
                  std::to_string(count);
            return SHK_ERR_ARG;
        }
        IOCHK(hipEventSynchronize(s->ev_copied));
        memcpy(host, s->host + s->off[j], count * sizeof(double));
        return SHK_OK;
    }
    err = "field " + std::to_string(id) + " is not in the last snapshot";
    return SHK_ERR_ARG;
}

namespace {

bool write_all(FILE* f, const void* p, size_t bytes) { return fwrite(p, 1, bytes, f) == bytes; }

// The file write proper; runs on the caller's thread or the writer thread.
int write_file(Snapshot* s, std::string path, int nx, int ny, double lx, double ly, int64_t t, double total_time,
               std::string& err) {
    hipError_t e = hipEventSynchronize(s->ev_copied);
    if (e != hipSuccess) {
        err = std::string("hipEventSynchronize: ") + hipGetErrorString(e);
        return SHK_ERR_HIP;
    }
    const int nf = (int)s->ids.size();
    Header h{};
    memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.nfields = (uint32_t)nf;
    h.nx = nx;
    h.ny = ny;
    h.t = t;
    h.total_time = total_time;
    h.lx = lx;
    h.ly = ly;
    const size_t table_end = sizeof(Header) + (size_t)nf * sizeof(Entry);
    h.data_offset = (table_end + 511) / 512 * 512;
    std::vector<Entry> ent(nf);
    uint64_t pos = h.data_offset;
    for (int j = 0; j < nf; ++j) {
        const FieldDesc& d = s->desc[s->ids[j]];
        Entry& E = ent[j];
        memset(&E, 0, sizeof(E));
        strncpy(E.name, d.name, sizeof(E.name) - 1);
        E.field_id = s->ids[j];
        E.rows = d.rows;
        E.cols = d.cols;
        E.offset = pos;
        E.bytes = s->cnt[j] * sizeof(double);
        E.checksum = s->host_ck[j];
        pos += (E.bytes + 511) / 512 * 512;
    }
    const std::string tmp = path + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot open " + tmp + ": " + strerror(errno);
        return SHK_ERR_ARG;
    }
    static const char zeros[512] = {0};
    bool ok = write_all(f, &h, sizeof(h)) && write_all(f, ent.data(), nf * sizeof(Entry)) &&
              write_all(f, zeros, h.data_offset - table_end);
    for (int j = 0; ok && j < nf; ++j) {
        ok = write_all(f, s->host + s->off[j], ent[j].bytes);
        const size_t pad = (512 - ent[j].bytes % 512) % 512;
        if (ok && pad) ok = write_all(f, zeros, pad);
    }
    ok = ok && fflush(f) == 0 && fsync(fileno(f)) == 0;
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        err = "writing " + tmp + " failed: " + strerror(errno);
        unlink(tmp.c_str());
        return SHK_ERR_ARG;
    }
    if (rename(tmp.c_str(), path.c_str()) != 0) {
        err = "cannot rename " + tmp + " to " + path + ": " + strerror(errno);
        unlink(tmp.c_str());
        return SHK_ERR_ARG;
    }
    return SHK_OK;
}

}  // namespace

int write_checkpoint(Snapshot* s, const char* path, int nx, int ny, double lx, double ly, int64_t t,
                     double total_time, bool async, std::string& err) {
    if (!s->have || (int)s->ids.size() != kMaxFields) {
        err = "a checkpoint needs a snapshot of every State field";
        return SHK_ERR_STATE;
    }
    join_writer(s);
    if (s->writer_rc != SHK_OK) {  // report a failed earlier write before starting another
        const int rc = s->writer_rc;
        err = "the previous checkpoint write failed: " + s->writer_err;
        s->writer_rc = SHK_OK;
        s->writer_err.clear();
        return rc;
    }
    std::string p(path);
    if (!async) return write_file(s, p, nx, ny, lx, ly, t, total_time, err);
    s->writer_rc = SHK_OK;
    s->writer_err.clear();
    s->writing = true;
    s->writer = std::thread([s, p, nx, ny, lx, ly, t, total_time]() {
        std::string e;
        const int rc = write_file(s, p, nx, ny, lx, ly, t, total_time, e);
        s->writer_rc = rc;
        s->writer_err = e;
    });
    return SHK_OK;
}

int wait(Snapshot* s, std::string& err) {
    join_writer(s);
    const int rc = s->writer_rc;
    if (rc != SHK_OK) err = s->writer_err;
    s->writer_rc = SHK_OK;
    s->writer_err.clear();
    return rc;
}

int read_checkpoint(Snapshot* s, hipStream_t st, const char* path, int nx, int ny, int nfields,
                    const FieldDesc* all, int64_t* t, double* total_time, std::string& err) {
    join_writer(s);
    FILE* f = fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open ") + path + ": " + strerror(errno);
        return SHK_ERR_ARG;
    }
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{f};
    struct stat sb;
    if (fstat(fileno(f), &sb) != 0) {
        err = std::string("cannot stat ") + path;
        return SHK_ERR_ARG;
    }
    const uint64_t fsize = (uint64_t)sb.st_size;
    Header h;
    if (fread(&h, 1, sizeof(h), f) != sizeof(h) || memcmp(h.magic, kMagic, 8) != 0) {
        err = std::string(path) + " is not a checkpoint of this library";
        return SHK_ERR_ARG;
    }
    if (h.version != kVersion) {
        err = std::string(path) + ": checkpoint version " + std::to_string(h.version) + ", this library reads " +
              std::to_string(kVersion);
        return SHK_ERR_ARG;
    }
    if (h.nx != nx || h.ny != ny) {
        err = std::string(path) + " holds a " + std::to_string(h.nx) + " x " + std::to_string(h.ny) +
              " grid, the State is " + std::to_string(nx) + " x " + std::to_string(ny);
        return SHK_ERR_ARG;
    }
    if (h.nfields < 1 || h.nfields > 4096) {
        err = std::string(path) + ": bad field count";
        return SHK_ERR_ARG;
    }
    std::vector<Entry> ent(h.nfields);
    if (fread(ent.data(), sizeof(Entry), h.nfields, f) != h.nfields) {
        err = std::string(path) + " is truncated (field table)";
        return SHK_ERR_ARG;
    }
    // match every State field by name
    std::vector<int> which(nfields, -1);
    std::vector<size_t> off(nfields), cnt(nfields);
    size_t total = 0;
    for (int j = 0; j < nfields; ++j) {
        for (uint32_t k = 0; k < h.nfields; ++k)
            if (strncmp(ent[k].name, all[j].name, sizeof(ent[k].name)) == 0) which[j] = (int)k;
        if (which[j] < 0) {
            err = std::string(path) + " has no field " + all[j].name;
            return SHK_ERR_ARG;
        }
        const Entry& E = ent[which[j]];
        cnt[j] = (size_t)all[j].rows * (size_t)all[j].cols;
        if (E.rows != all[j].rows || E.cols != all[j].cols || E.bytes != cnt[j] * sizeof(double)) {
            err = std::string(path) + ": field " + all[j].name + " has the wrong shape";
            return SHK_ERR_ARG;
        }
        if (E.offset > fsize || E.bytes > fsize - E.offset) {
            err = std::string(path) + " is truncated (field " + all[j].name + ")";
            return SHK_ERR_ARG;
        }
        off[j] = total;
        total += align_up(cnt[j]);
    }
    if (int rc = reserve(s, st, total, err)) return rc;
    if (s->copied) IOCHK(hipEventSynchronize(s->ev_copied));  // the last snapshot's copy writes the host area
    s->have = false;
    for (int j = 0; j < nfields; ++j) {
        const Entry& E = ent[which[j]];
        if (fseeko(f, (off_t)E.offset, SEEK_SET) != 0 || fread(s->host + off[j], 1, E.bytes, f) != E.bytes) {
            err = std::string(path) + ": reading field " + all[j].name + " failed";
            return SHK_ERR_ARG;
        }
    }
    IOCHK(hipMemcpyAsync(s->dev, s->host, total * sizeof(double), hipMemcpyHostToDevice, st));
    for (int j = 0; j < nfields; ++j) checksum(s, st, s->dev + off[j], cnt[j], j);
    IOCHK(hipGetLastError());
    IOCHK(hipMemcpyAsync(s->host_ck, s->dev_ck, nfields * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    IOCHK(hipStreamSynchronize(st));
    for (int j = 0; j < nfields; ++j)
        if (s->host_ck[j] != ent[which[j]].checksum) {
            err = std::string(path) + ": field " + all[j].name + " fails its checksum; the State is unchanged";
            return SHK_ERR_NUMERIC;
        }
    for (int j = 0; j < nfields; ++j)
        IOCHK(hipMemcpyAsync(all[j].dev, s->dev + off[j], cnt[j] * sizeof(double), hipMemcpyDeviceToDevice, st));
    IOCHK(hipStreamSynchronize(st));
    *t = h.t;
    *total_time = h.total_time;
    return SHK_OK;
}

int field_stats(Snapshot* s, hipStream_t st, const double* dev, size_t n, double out[3], std::string& err) {
    if (n == 0) {
        err = "empty field";
        return SHK_ERR_ARG;
    }
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k_stats_part, dim3(nb), dim3(kThreads), 0, st, dev, n, s->stat_part);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kThreads), 0, st, s->stat_part, nb, (double)n, s->stat_dev);
    IOCHK(hipGetLastError());
    IOCHK(hipMemcpyAsync(s->stat_host, s->stat_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    IOCHK(hipStreamSynchronize(st));
    out[0] = s->stat_host[0];
    out[1] = s->stat_host[1];
    out[2] = s->stat_host[2];
    return SHK_OK;
}

}  // namespace shkio
