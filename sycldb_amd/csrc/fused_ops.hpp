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

// OR of a value over the active lanes of the wavefront (ROCm device library)
extern "C" __device__ __attribute__((const)) unsigned int __ockl_wfred_or_u32(unsigned int);

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

// N row values held as one vector value: element b (b per lane, not
// necessarily uniform) is then read with a register extract -- an indexed
// access to an array would put the array in scratch memory.
template <int N> struct Lanes {
    typedef int32_t type __attribute__((ext_vector_type(N)));
};

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

// Two-phase probes.  *_issue reads, for every flagged row whose key is in
// range, what the probe needs from the table (the presence word or table
// entry and, for value tables, the value, loaded together rather than one
// after the other); *_finish forms the result from it.  Code between the two
// phases (the next tile's column loads) is issued before the probe's loads
// are waited for.
template <int R> struct ProbeTmp {
    uint32_t word[R];
    int32_t val[R];
};

// A semi or full probe through the presence bitmap of the reference table
// (bit set = the table's match flag; FilterJoinKernel / FullJoinKernel,
// kernels/join.hpp:129-160, 256-300), with the matched value, when wanted,
// from a 16-bit offset table (FV_U16), the reference pairs (FV_PAIR) or a
// rank dictionary in LDS (FV_RANK16: `bits` is the dictionary, `words` its
// word count, and the issue phase reads the word's prefix count).
//
// The key range test mn <= x <= mx is the single unsigned test
// (x - mn) <= (mx - mn), exact whenever mx >= mn (an empty range matches
// nothing).  A bitmap in LDS (IN_LDS) is read for every row at an index
// clamped into the bitmap, so the read needs no branch; a bitmap in global
// memory is read only for the rows that need it.
template <int R, int VKIND, bool IN_LDS = false>
__device__ __forceinline__ void probe_bits_issue(const int32_t (&x)[R], uint32_t in, const uint32_t *bits, int32_t mn,
                                                 int32_t mx, const void *vals, int32_t vmin, int32_t words,
                                                 ProbeTmp<R> &t)
{
    const uint32_t span = (uint32_t)mx - (uint32_t)mn;
    const bool nonempty = mx >= mn;
#pragma unroll
    for (int k = 0; k < R; k++) {
        t.word[k] = 0;
        t.val[k] = 0;
        const uint32_t s = (uint32_t)x[k] - (uint32_t)mn;
        if constexpr (IN_LDS && VKIND == FV_NONE) {
            if (nonempty) t.word[k] = bits[(s < span ? s : span) >> 5];
        } else {
            const bool go = nonempty & (((in >> k) & 1u) != 0) & (s <= span);
            if (go) {
                t.word[k] = bits[s >> 5];
                if constexpr (VKIND == FV_U16)
                    t.val[k] = (int32_t)((uint32_t)vmin + reinterpret_cast<const uint16_t *>(vals)[s]);
                else if constexpr (VKIND == FV_PAIR)
                    t.val[k] = reinterpret_cast<const int2 *>(vals)[s].y;
                else if constexpr (VKIND == FV_RANK16)
                    t.val[k] = reinterpret_cast<const uint16_t *>(reinterpret_cast<const uint8_t *>(bits) +
                                                                  rank_prefix_off((uint32_t)words))[s >> 5];
            }
        }
    }
}

template <int R, int VKIND, bool IN_LDS = false>
__device__ __forceinline__ uint32_t probe_bits_finish(const int32_t (&x)[R], uint32_t in, int32_t mn, int32_t mx,
                                                      const ProbeTmp<R> &t, const uint32_t *bits, int32_t vmin,
                                                      int32_t words, int32_t (&v)[R])
{
    const uint32_t span = (uint32_t)mx - (uint32_t)mn;
    const bool nonempty = mx >= mn;
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const uint32_t s = (uint32_t)x[k] - (uint32_t)mn;
        const bool hit = nonempty & (((in >> k) & 1u) != 0) & (s <= span) & (((t.word[k] >> (s & 31)) & 1u) != 0);
        out |= (uint32_t)hit << k;
        if constexpr (VKIND == FV_RANK16) {
            if (hit) {
                const uint32_t rank = (uint32_t)t.val[k] + __popc(t.word[k] & ((1u << (s & 31)) - 1u));
                v[k] = (int32_t)((uint32_t)vmin + reinterpret_cast<const uint16_t *>(
                                                      reinterpret_cast<const uint8_t *>(bits) +
                                                      rank_values_off((uint32_t)words))[rank]);
            }
        } else if constexpr (VKIND != FV_NONE) {
            v[k] = hit ? t.val[k] : v[k];
        }
    }
    return out;
}

// A bitmap probe answered by its FV_U16 value table alone: a slot reads
// VALS_UNSET exactly when its presence bit is clear, so one read gives both
// the match and the value (no bitmap read).  Rows outside the key range or
// the input flags read nothing and keep VALS_UNSET.
template <int R>
__device__ __forceinline__ void probe_unset_issue(const int32_t (&x)[R], uint32_t in, const uint16_t *vals,
                                                  int32_t mn, int32_t mx, ProbeTmp<R> &t)
{
    const uint32_t span = (uint32_t)mx - (uint32_t)mn;
    const bool nonempty = mx >= mn;
#pragma unroll
    for (int k = 0; k < R; k++) {
        t.word[k] = VALS_UNSET;
        const uint32_t s = (uint32_t)x[k] - (uint32_t)mn;
        if (nonempty & (((in >> k) & 1u) != 0) & (s <= span)) t.word[k] = vals[s];
    }
}

template <int R>
__device__ __forceinline__ uint32_t probe_unset_finish(uint32_t in, const ProbeTmp<R> &t, int32_t vmin,
                                                       int32_t (&v)[R])
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const bool hit = (((in >> k) & 1u) != 0) & (t.word[k] != VALS_UNSET);
        out |= (uint32_t)hit << k;
        v[k] = hit ? (int32_t)((uint32_t)vmin + t.word[k]) : v[k];
    }
    return out;
}

template <int R>
__device__ __forceinline__ void probe_semi_issue(const int32_t (&x)[R], uint32_t in, const uint8_t *ht, int32_t mn,
                                                 int32_t mx, ProbeTmp<R> &t)
{
    const uint32_t span = (uint32_t)mx - (uint32_t)mn;
    const bool nonempty = mx >= mn;
#pragma unroll
    for (int k = 0; k < R; k++) {
        t.word[k] = 0;
        const uint32_t s = (uint32_t)x[k] - (uint32_t)mn;
        if (nonempty & (((in >> k) & 1u) != 0) & (s <= span)) t.word[k] = ht[s] != 0;
    }
}

template <int R>
__device__ __forceinline__ uint32_t probe_semi_finish(uint32_t in, const ProbeTmp<R> &t)
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++) out |= (uint32_t)((((in >> k) & 1u) != 0) & (t.word[k] != 0)) << k;
    return out;
}

template <int R>
__device__ __forceinline__ void probe_full_issue(const int32_t (&x)[R], uint32_t in, const int2 *ht, int32_t mn,
                                                 int32_t mx, ProbeTmp<R> &t)
{
    const uint32_t span = (uint32_t)mx - (uint32_t)mn;
    const bool nonempty = mx >= mn;
#pragma unroll
    for (int k = 0; k < R; k++) {
        t.word[k] = 0;
        t.val[k] = 0;
        const uint32_t s = (uint32_t)x[k] - (uint32_t)mn;
        if (nonempty & (((in >> k) & 1u) != 0) & (s <= span)) {
            const int2 e = ht[s];
            t.word[k] = e.x == 1;
            t.val[k] = e.y;
        }
    }
}

template <int R>
__device__ __forceinline__ uint32_t probe_full_finish(uint32_t in, const ProbeTmp<R> &t, int32_t (&v)[R])
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const bool hit = (((in >> k) & 1u) != 0) & (t.word[k] != 0);
        out |= (uint32_t)hit << k;
        v[k] = hit ? t.val[k] : v[k];
    }
    return out;
}

// The value of a full probe whose match the shared masks of a batched launch
// have already established for the rows that matter: the table's value for
// every row of `in` whose key is in range (FV_U16, FV_PAIR, or the reference
// pairs of a FullJoinKernel table when VKIND is FV_NONE).
template <int R, int VKIND>
__device__ __forceinline__ void probe_value_issue(const int32_t (&x)[R], uint32_t in, const void *vals, int32_t mn,
                                                  int32_t mx, int32_t vmin, ProbeTmp<R> &t)
{
#pragma unroll
    for (int k = 0; k < R; k++) {
        t.val[k] = 0;
        if (((in >> k) & 1u) && x[k] >= mn && x[k] <= mx) {
            const uint32_t s = slot_of(x[k], mn);
            if constexpr (VKIND == FV_U16)
                t.val[k] = (int32_t)((uint32_t)vmin + reinterpret_cast<const uint16_t *>(vals)[s]);
            else
                t.val[k] = reinterpret_cast<const int2 *>(vals)[s].y;
        }
    }
}

// The value of a probe's matched rows `out`, fetched after the match (the
// issue phase then reads only the presence words).
template <int R, int VKIND>
__device__ __forceinline__ void probe_value_after(const int32_t (&x)[R], uint32_t out, const void *vals, int32_t mn,
                                                  int32_t vmin, int32_t (&v)[R])
{
#pragma unroll
    for (int k = 0; k < R; k++)
        if ((out >> k) & 1u) {
            const uint32_t s = slot_of(x[k], mn);
            if constexpr (VKIND == FV_U16)
                v[k] = (int32_t)((uint32_t)vmin + reinterpret_cast<const uint16_t *>(vals)[s]);
            else
                v[k] = reinterpret_cast<const int2 *>(vals)[s].y;
        }
}

template <int R> __device__ __forceinline__ void probe_value_finish(uint32_t out, const ProbeTmp<R> &t, int32_t (&v)[R])
{
#pragma unroll
    for (int k = 0; k < R; k++)
        if ((out >> k) & 1u) v[k] = t.val[k];
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
