// Device pieces shared by the fused bundle kernels: the compiled program
// interpreter (fused.hip) and the kernels compiled at run time for one
// program shape (jit.cpp, which sees this file as an in-memory header).
//
// Flag registers of a chunk of R rows are R-bit masks; a buffer of 1-byte
// flags is read and written 4 rows at a time.  The group-by row update and
// the block flush implement GroupByAggregateKernel
// (kernels/aggregation.hpp:278-409) with LDS-privatised partial sums.
#pragma once

#include "fused_types.hpp"
#include "sdb_semantics.hpp"

namespace sdb {
namespace fdev {

constexpr int BLOCK = 256;
constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;

typedef int v4i __attribute__((ext_vector_type(4)));

template <int R> __device__ __forceinline__ uint32_t flags_in(const uint8_t *p)
{
    if constexpr (R == 4) {
        const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
        return ((w & 0xffu) != 0) | (((w >> 8) & 0xffu) != 0) << 1 | (((w >> 16) & 0xffu) != 0) << 2 |
               (((w >> 24) & 0xffu) != 0) << 3;
    } else {
        return p[0] != 0;
    }
}

template <int R> __device__ __forceinline__ void flags_out(uint8_t *p, uint32_t m)
{
    if constexpr (R == 4) {
        *reinterpret_cast<uint32_t *>(p) = (m & 1u) | ((m >> 1) & 1u) << 8 | ((m >> 2) & 1u) << 16 | ((m >> 3) & 1u) << 24;
    } else {
        p[0] = (uint8_t)(m & 1u);
    }
}

__device__ __forceinline__ uint32_t logic_combine(uint32_t logic, uint32_t in, uint32_t pred)
{
    return logic == SDB_NONE ? pred : (logic == SDB_AND ? (in & pred) : (in | pred));
}

// Adds a block's partial sum of slot s to the global result; the first block
// to reach a slot writes its key values (first key varies fastest), which is
// exact for in-range keys.
__device__ __forceinline__ void gb_flush_slot(const FGroup &g, uint32_t s, unsigned long long sum)
{
    atomicAdd(reinterpret_cast<unsigned long long *>(&g.sums[s]), sum);
    if (atomicExch(&g.seen[s], 1u) == 0u) {
        uint32_t rest = s;
#pragma unroll
        for (int j = 0; j < F_MAX_KEYS; j++) {
            if (j >= g.k) break;
            g.out_keys[j][s] = (int32_t)((uint32_t)g.mins[j] + rest % g.range[j]);
            rest /= g.range[j];
        }
    }
}

// One flagged row: slot = (sum_j (key_j - min_j) * prod_{l<j} range_l) as
// int32, then % prod in uint64.  Rows whose keys are all in range go to the
// block's LDS copy (dense, or an open-addressed hash of slots); others, and
// rows meeting a crowded hash, update the global arrays directly.  MODE is
// the FGroupMode (fixed per program, so only its path is compiled).
template <int MODE>
__device__ __forceinline__ void gb_row_mode(const FGroup &g, uint8_t *lds, const int32_t *kv, int32_t agg)
{
    bool in_range = true;
    uint32_t hash = 0, mult = 1;
#pragma unroll
    for (int j = 0; j < F_MAX_KEYS; j++) {
        if (j >= g.k) break;
        in_range &= kv[j] >= g.mins[j] && kv[j] <= g.maxs[j];
        hash += ((uint32_t)kv[j] - (uint32_t)g.mins[j]) * mult;
        mult *= g.range[j];
    }
    const unsigned long long v = (unsigned long long)(int64_t)agg;
    // LDS modes are chosen only when the key ranges multiply to exactly prod
    // < 2^32, so an in-range row's slot is `hash` itself (no modulo needed)
    if constexpr (MODE == FG_DENSE_LDS) {
        if (in_range) {
            const uint32_t s = hash;
            unsigned long long *lsum = reinterpret_cast<unsigned long long *>(lds + g.lds_off);
            uint32_t *lbits = reinterpret_cast<uint32_t *>(lsum + g.prod);
            atomicAdd(&lsum[s], v);
            atomicOr(&lbits[s >> 5], 1u << (s & 31));
            return;
        }
    }
    if constexpr (MODE == FG_HASH_LDS) {
        if (in_range) {
            const uint32_t s = hash;
            uint32_t *hk = reinterpret_cast<uint32_t *>(lds + g.lds_off);
            unsigned long long *hv = reinterpret_cast<unsigned long long *>(hk + g.hash_cap);
            const uint32_t mask = g.hash_cap - 1;
            uint32_t h = (s * 2654435761u) & mask;
#pragma unroll 1
            for (int probe = 0; probe < 32; probe++) {
                uint32_t cur = hk[h];
                if (cur == HASH_EMPTY) {
                    cur = atomicCAS(&hk[h], HASH_EMPTY, s);
                    if (cur == HASH_EMPTY) cur = s;
                }
                if (cur == s) {
                    atomicAdd(&hv[h], v);
                    return;
                }
                h = (h + 1) & mask;
            }
            // table crowded: fall through to the global update
        }
    }
    const uint64_t s64 = (uint64_t)(int64_t)(int32_t)hash % g.prod;
    // a slot already marked needs no exchange (a plain read; a stale 0 just
    // leads to the exchange, which decides)
    if (__hip_atomic_load(&g.seen[s64], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u &&
        atomicExch(&g.seen[s64], 1u) == 0u) {
#pragma unroll
        for (int j = 0; j < F_MAX_KEYS; j++)
            if (j < g.k) g.out_keys[j][s64] = kv[j];
    }
    atomicAdd(reinterpret_cast<unsigned long long *>(&g.sums[s64]), v);
}

// The same with the mode read from the program.
__device__ __forceinline__ void gb_row(const FGroup &g, uint8_t *lds, const int32_t *kv, int32_t agg)
{
    if (g.lds == FG_DENSE_LDS) gb_row_mode<FG_DENSE_LDS>(g, lds, kv, agg);
    else if (g.lds == FG_HASH_LDS) gb_row_mode<FG_HASH_LDS>(g, lds, kv, agg);
    else gb_row_mode<FG_GLOBAL>(g, lds, kv, agg);
}

// Block start: stage probe bitmaps into LDS and clear the aggregation area
// (NT: threads per block).
template <int NT = BLOCK> __device__ __forceinline__ void block_prologue(const FProgram &P, uint8_t *lds)
{
    for (int b = 0; b < P.nblobs; b++) {
        const uint4 *src = reinterpret_cast<const uint4 *>(P.blob_src[b]);
        uint4 *dst = reinterpret_cast<uint4 *>(lds + P.blob_off[b]);
        for (uint32_t w = threadIdx.x; w < P.blob_bytes[b] / 16; w += NT) dst[w] = src[w];
    }
    const FGroup &g = P.gb;
    if (g.lds == FG_DENSE_LDS) {
        unsigned long long *lsum = reinterpret_cast<unsigned long long *>(lds + g.lds_off);
        uint32_t *lbits = reinterpret_cast<uint32_t *>(lsum + g.prod);
        for (uint32_t s = threadIdx.x; s < (uint32_t)g.prod; s += NT) lsum[s] = 0;
        for (uint32_t w = threadIdx.x; w < ((uint32_t)g.prod + 31) / 32; w += NT) lbits[w] = 0;
    } else if (g.lds == FG_HASH_LDS) {
        uint32_t *hk = reinterpret_cast<uint32_t *>(lds + g.lds_off);
        unsigned long long *hv = reinterpret_cast<unsigned long long *>(hk + g.hash_cap);
        for (uint32_t h = threadIdx.x; h < g.hash_cap; h += NT) {
            hk[h] = HASH_EMPTY;
            hv[h] = 0;
        }
    }
    __syncthreads();
}

// Block end: the terminal sum (one atomic per block) and the flush of the
// block's LDS group-by partials (NT: threads per block).
template <int NT = BLOCK> __device__ __forceinline__ void block_epilogue(const FProgram &P, uint8_t *lds, int64_t acc)
{
    if (P.sum_acc) {
        __shared__ int64_t part[NT / 64];
        uint64_t v = (uint64_t)acc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = (int64_t)v;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t s = 0;
            for (int w = 0; w < NT / 64; w++) s += (uint64_t)part[w];
            if (s) atomicAdd(reinterpret_cast<unsigned long long *>(P.sum_acc), (unsigned long long)s);
        }
    }
    const FGroup &g = P.gb;
    if (g.lds == FG_DENSE_LDS) {
        __syncthreads();
        const unsigned long long *lsum = reinterpret_cast<const unsigned long long *>(lds + g.lds_off);
        const uint32_t *lbits = reinterpret_cast<const uint32_t *>(lsum + g.prod);
        for (uint32_t s = threadIdx.x; s < (uint32_t)g.prod; s += NT)
            if ((lbits[s >> 5] >> (s & 31)) & 1u) gb_flush_slot(g, s, lsum[s]);
    } else if (g.lds == FG_HASH_LDS) {
        __syncthreads();
        const uint32_t *hk = reinterpret_cast<const uint32_t *>(lds + g.lds_off);
        const unsigned long long *hv = reinterpret_cast<const unsigned long long *>(hk + g.hash_cap);
        for (uint32_t h = threadIdx.x; h < g.hash_cap; h += NT)
            if (hk[h] != HASH_EMPTY) gb_flush_slot(g, hk[h], hv[h]);
    }
}

} // namespace fdev
} // namespace sdb
