// Shared definitions of the SYCLDB operator library for MI355X (gfx950).
//
// The per-row integer semantics live in sdb_semantics.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "sdb_semantics.hpp"

namespace sdb {

// Error carrying a C-ABI status code; converted at the boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        fail(SDB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define SDB_HIP(call) ::sdb::hip_check((call), #call)

// ------------------------------------------------------------- launch shape

// Number of compute units of the current device (cached per device).
int device_cu_count();
// Grid for a grid-stride streaming kernel over `units` work items with
// `per_block` items consumed per block iteration.
unsigned stream_grid(int64_t units, int64_t per_block, int blocks_per_cu = 8);

// Library-owned scratch: a device buffer of at least `bytes`, reused across
// calls on the same device and stream-ordered by the caller's stream.
void *scratch_buffer(size_t bytes, hipStream_t stream);

} // namespace sdb
