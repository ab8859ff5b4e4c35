// Checkpoint / restart and field output of the device State (shk_io.hip).
// Internal to libshakti_hip.so; the entry points are in shk_api.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace shkio {

struct Snapshot;

// One per context, created on first use.  Owns a copy stream (default
// priority), a device staging area, a pinned host staging area and the
// checkpoint writer thread.
Snapshot* create(std::string& err);
void destroy(Snapshot* s);  // joins the writer, frees everything

// The face/cell shape of every field, by id, and its name in the reference's
// State (src/state.jl:13-57).
struct FieldDesc {
    const char* name;
    int rows, cols;
    double* dev;
};

// Copies the listed fields into the device staging area on `st` (the State
// can change as soon as this returns), computes each one's checksum there and
// queues the device-to-host copy on the copy stream.  Waits for a running
// checkpoint writer first.
int begin(Snapshot* s, hipStream_t st, int nfields, const int* ids, const FieldDesc* all, std::string& err);

// Waits for the copy of the last snapshot and copies field `id` out of it.
int fetch(Snapshot* s, int id, double* host, size_t count, std::string& err);

// Writes the last snapshot (which must hold every field) to `path`: into
// path + ".tmp", flushed to disk, then renamed over `path`.  With async the
// write runs on a host thread and `wait` reports its outcome.
int write_checkpoint(Snapshot* s, const char* path, int nx, int ny, double lx, double ly, int64_t t,
                     double total_time, bool async, std::string& err);
int wait(Snapshot* s, std::string& err);

// Reads a checkpoint, checks its grid, field table and every checksum on the
// device, and only then copies the fields into the State (all or nothing).
int read_checkpoint(Snapshot* s, hipStream_t st, const char* path, int nx, int ny, int nfields,
                    const FieldDesc* all, int64_t* t, double* total_time, std::string& err);

// min, max and mean of one device array (NaN anywhere gives NaN for all three
// statistics, as extrema and mean do).
int field_stats(Snapshot* s, hipStream_t st, const double* dev, size_t n, double out[3], std::string& err);

}  // namespace shkio
