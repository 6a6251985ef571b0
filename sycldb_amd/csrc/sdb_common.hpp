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

// ------------------------------------------------------- result transfer

// Ids of the rows whose flag is set, in row order, into row_ids (capacity n);
// their number into *count.  tile_counts: compact_tiles(n) words of device
// memory for the per-tile counts.
int64_t compact_tiles(int64_t n);
void launch_compact_into(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, uint32_t *tile_counts,
                         hipStream_t s);
// dst[c][j] = src[c][ids[j]] for j < count (0 where ids[j] >= src_rows[c]);
// a column holds int32 values, or uint64 where bit c of is_u64 is set.
constexpr int GATHER_MAX_COLS = 8;
struct GatherCols {
    const void *src[GATHER_MAX_COLS];
    void *dst[GATHER_MAX_COLS];
    uint64_t src_rows[GATHER_MAX_COLS];
    uint32_t is_u64;
    int32_t ncols;
};
void launch_gather_cols(const GatherCols &g, const int32_t *ids, int64_t count, hipStream_t s);
// Copies device buffers src[i] (bytes[i] bytes) to dst + off[i]: several
// small arrays packed into one buffer for a single transfer.
constexpr int PACK_MAX_PARTS = 64;
struct PackParts {
    const uint8_t *src[PACK_MAX_PARTS];
    uint64_t off[PACK_MAX_PARTS];
    uint64_t bytes[PACK_MAX_PARTS];
    int32_t n;
};
void launch_pack_parts(const PackParts &p, uint8_t *dst, hipStream_t s);

// Statistics of compact probe tables, each stats[i] laid out as [overflow
// flag, nparts[i] partial counts]: the sums and flags go to dst as pairs.
struct StatsParts {
    const uint32_t *stats[PACK_MAX_PARTS];
    uint32_t nparts[PACK_MAX_PARTS];
    int32_t n;
};
void launch_stats_reduce(const StatsParts &p, uint32_t *dst, hipStream_t s);

} // namespace sdb
