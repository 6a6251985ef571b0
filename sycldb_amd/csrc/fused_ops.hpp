// Per-instruction device templates for the fused kernels compiled at run time
// for one program shape (jit.cpp generates the code that calls them; this
// file reaches the compiler as an in-memory header).
//
// Each template works on one chunk of R rows (R = 4: a 16-byte slice of each
// column; R = 1: a single row) and reproduces the per-row result of the
// reference functor named beside it, exactly as the interpreter in fused.hip
// does; flag arguments are R-bit masks.  Instruction parameters that only
// select code (comparison, logical and arithmetic op, table placement) are
// template arguments, so the generated kernel has no dispatch left.
#pragma once

#include "fused_device.hpp"

namespace sdb {
namespace fops {

using fdev::v4i;

template <int R> struct Vec { typedef v4i type; };
template <> struct Vec<1> { typedef int32_t type; };

template <int R> __device__ __forceinline__ typename Vec<R>::type load_vec(const int32_t *mem, int64_t row0)
{
    if constexpr (R == 4) return __builtin_nontemporal_load(reinterpret_cast<const v4i *>(mem + row0));
    else return __builtin_nontemporal_load(mem + row0);
}

template <int R> __device__ __forceinline__ void vec_rows(const typename Vec<R>::type &v, int32_t (&x)[R])
{
    if constexpr (R == 4) {
        x[0] = v.x;
        x[1] = v.y;
        x[2] = v.z;
        x[3] = v.w;
    } else {
        x[0] = v;
    }
}

// R rows of a column, streamed (non-temporal)
template <int R> __device__ __forceinline__ void load_rows(const int32_t *mem, int64_t row0, int32_t (&x)[R])
{
    vec_rows<R>(load_vec<R>(mem, row0), x);
}

template <int R> __device__ __forceinline__ void copy_rows(const int32_t (&src)[R], int32_t (&x)[R])
{
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = src[k];
}

template <int R> __device__ __forceinline__ void fill_rows(int32_t v, int32_t (&x)[R])
{
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = v;
}

// masked register write: rows with a set bit take the new value
template <int R> __device__ __forceinline__ void set_rows(int32_t (&dst)[R], const int32_t (&src)[R], uint32_t m)
{
#pragma unroll
    for (int k = 0; k < R; k++)
        if ((m >> k) & 1u) dst[k] = src[k];
}

template <int R> __device__ __forceinline__ void store_rows(int32_t *dst, int64_t row0, const int32_t (&x)[R], uint32_t m)
{
    constexpr uint32_t M = (1u << R) - 1;
    if constexpr (R == 4) {
        if (m == M) {
            *reinterpret_cast<int4 *>(dst + row0) = make_int4(x[0], x[1], x[2], x[3]);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < R; k++)
        if ((m >> k) & 1u) dst[row0 + k] = x[k];
}

// rows a selection with logical op LOGIC has to evaluate
template <int LOGIC, int R> __device__ __forceinline__ uint32_t sel_need(uint32_t in)
{
    constexpr uint32_t M = (1u << R) - 1;
    return LOGIC == SDB_NONE ? M : (LOGIC == SDB_AND ? in : (~in & M));
}

template <int LOGIC> __device__ __forceinline__ uint32_t combine(uint32_t in, uint32_t pred)
{
    return LOGIC == SDB_NONE ? pred : (LOGIC == SDB_AND ? (in & pred) : (in | pred));
}

// SelectionKernelLiteral (kernels/selection.hpp:98-134): cmp(x, lit)
template <int R, int CMP> __device__ __forceinline__ uint32_t pred_lit(const int32_t (&x)[R], int32_t lit)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < R; k++) m |= (uint32_t)compare(CMP, x[k], lit) << k;
    return m;
}

// SelectionKernelColumns (kernels/selection.hpp:136-168): cmp(x, y)
template <int R, int CMP> __device__ __forceinline__ uint32_t pred_col(const int32_t (&x)[R], const int32_t (&y)[R])
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < R; k++) m |= (uint32_t)compare(CMP, x[k], y[k]) << k;
    return m;
}

// SEARCH over one range (models/models.hpp:604-718): lo <= x <= hi
template <int R> __device__ __forceinline__ uint32_t pred_range(const int32_t (&x)[R], int32_t lo, int32_t hi)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < R; k++) m |= (uint32_t)(x[k] >= lo && x[k] <= hi) << k;
    return m;
}

// SEARCH over two points: x == a || x == b
template <int R> __device__ __forceinline__ uint32_t pred_in2(const int32_t (&x)[R], int32_t a, int32_t b)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < R; k++) m |= (uint32_t)(x[k] == a || x[k] == b) << k;
    return m;
}

// PerformOperationKernel* (kernels/aggregation.hpp:37-130) on the rows of `in`
template <int R, int OP>
__device__ __forceinline__ void binop(const int32_t (&x)[R], const int32_t (&y)[R], int32_t (&r)[R])
{
#pragma unroll
    for (int k = 0; k < R; k++) r[k] = element_op(OP, x[k], y[k]);
}

// FilterJoinKernel (kernels/join.hpp:129-160): 1-byte table
template <int R>
__device__ __forceinline__ uint32_t probe_semi(const int32_t (&x)[R], uint32_t in, const uint8_t *ht, int32_t mn,
                                               int32_t mx)
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++)
        if (((in >> k) & 1u) && x[k] >= mn && x[k] <= mx && ht[slot_of(x[k], mn)] != 0) out |= 1u << k;
    return out;
}

// FullJoinKernel (kernels/join.hpp:256-300): (flag, value) pairs; matched
// rows take the value
template <int R>
__device__ __forceinline__ uint32_t probe_full(const int32_t (&x)[R], uint32_t in, const int2 *ht, int32_t mn,
                                               int32_t mx, int32_t (&v)[R])
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (((in >> k) & 1u) && x[k] >= mn && x[k] <= mx) {
            const int2 e = ht[slot_of(x[k], mn)];
            if (e.x == 1) {
                v[k] = e.y;
                out |= 1u << k;
            }
        }
    }
    return out;
}

// A semi or full probe through the presence bitmap of the reference table
// (bit set = the table's match flag), with the matched value, when wanted,
// from a 16-bit offset table (FV_U16) or the reference pairs (FV_PAIR).
template <int R, int VKIND>
__device__ __forceinline__ uint32_t probe_bits(const int32_t (&x)[R], uint32_t in, const uint32_t *bits,
                                               int32_t mn, int32_t mx, const void *vals, int32_t vmin,
                                               int32_t (&v)[R])
{
    uint32_t out = 0, slot[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
        slot[k] = slot_of(x[k], mn);
        if (((in >> k) & 1u) && x[k] >= mn && x[k] <= mx && ((bits[slot[k] >> 5] >> (slot[k] & 31)) & 1u))
            out |= 1u << k;
    }
    if constexpr (VKIND != FV_NONE) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            if ((out >> k) & 1u) {
                if constexpr (VKIND == FV_U16)
                    v[k] = (int32_t)((uint32_t)vmin + reinterpret_cast<const uint16_t *>(vals)[slot[k]]);
                else
                    v[k] = reinterpret_cast<const int2 *>(vals)[slot[k]].y;
            }
        }
    }
    return out;
}

// BuildKeysHTKernel (kernels/join.hpp:17-40): every row writes its flag
template <int R>
__device__ __forceinline__ void build_keys(const int32_t (&x)[R], uint32_t in, uint8_t *ht, int32_t mn, int32_t len)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        const uint32_t s = slot_of(x[k], mn);
        if (s < (uint32_t)len) ht[s] = (in >> k) & 1u;
    }
}

// BuildKeyValsHTKernel (kernels/join.hpp:60-90): (1, value) for flagged rows,
// flag 0 for the others
template <int R>
__device__ __forceinline__ void build_keyvals(const int32_t (&x)[R], const int32_t (&y)[R], uint32_t in, int32_t *ht,
                                              int32_t mn, int32_t len)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        const uint32_t s = slot_of(x[k], mn);
        if (s >= (uint32_t)len) continue;
        if ((in >> k) & 1u) reinterpret_cast<int2 *>(ht)[s] = make_int2(1, y[k]);
        else ht[2 * (int64_t)s] = 0;
    }
}

// AggregateOperationKernel (kernels/aggregation.hpp:208-276): sum of the
// flagged rows as int64 (the block reduction adds it to the uint64 result)
template <int R> __device__ __forceinline__ void sum_rows(const int32_t (&x)[R], uint32_t in, int64_t &acc)
{
#pragma unroll
    for (int k = 0; k < R; k++)
        if ((in >> k) & 1u) acc += x[k];
}

} // namespace fops
} // namespace sdb
