// Operator kernels of the drop-in per-operator path (one launch per
// reference functor) and their C-ABI entry points.
//
// Layout: int32 columns and 1-byte flags in HBM.  A thread handles chunks of
// 4 consecutive rows (one 16-byte column load, one 32-bit flag word), a block
// handles 256 x UNROLL chunks per step, the grid strides over the column.  All
// loads of a step are issued before any store so each wave keeps UNROLL
// vector loads in flight.  Comparison / logic / arithmetic variants are
// template parameters so the row loop carries no data-independent branches.
// Pointers that are not 16-byte (columns) or 4-byte (flags) aligned take a
// one-row-per-thread path instead.
#include "capi_util.hpp"
#include "fused_types.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace sdb {
namespace {

constexpr int BLOCK = 256;
constexpr int UNROLL = 4;
constexpr int64_t ROWS_PER_STEP = (int64_t)BLOCK * UNROLL * 4;

inline bool aligned(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int k) { return (w >> (8 * k)) & 0xffu; }

// ------------------------------------------------------------- selection

template <int CMP, int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_select_lit(const int32_t *__restrict__ col,
                                                      const uint8_t *in, uint8_t *out, int64_t n,
                                                      int32_t v)
{
    const int64_t nchunks = n >> 2;
    const int4 *c4 = reinterpret_cast<const int4 *>(col);
    const uint32_t *in32 = reinterpret_cast<const uint32_t *>(in);
    uint32_t *out32 = reinterpret_cast<uint32_t *>(out);
    for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
         base += (int64_t)gridDim.x * BLOCK * UNROLL) {
        int4 x[UNROLL];
        uint32_t f[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nchunks) {
                x[u] = c4[c];
                f[u] = (LOGIC == SDB_NONE) ? 0u : in32[c];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nchunks) {
                const int32_t xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
                uint32_t r = 0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    r |= (uint32_t)logical(LOGIC, byte_of(f[u], k) != 0, compare(CMP, xs[k], v)) << (8 * k);
                out32[c] = r;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
        const int64_t i = (nchunks << 2) + threadIdx.x;
        out[i] = logical(LOGIC, in[i] != 0, compare(CMP, col[i], v));
    }
}

template <int CMP, int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_select_lit_row(const int32_t *col, const uint8_t *in,
                                                          uint8_t *out, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = logical(LOGIC, in[i] != 0, compare(CMP, col[i], v));
}

template <int CMP, int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_select_col(const int32_t *__restrict__ a,
                                                      const int32_t *__restrict__ b,
                                                      const uint8_t *in, uint8_t *out, int64_t n)
{
    const int64_t nchunks = n >> 2;
    const int4 *a4 = reinterpret_cast<const int4 *>(a);
    const int4 *b4 = reinterpret_cast<const int4 *>(b);
    const uint32_t *in32 = reinterpret_cast<const uint32_t *>(in);
    uint32_t *out32 = reinterpret_cast<uint32_t *>(out);
    for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
         base += (int64_t)gridDim.x * BLOCK * UNROLL) {
        int4 x[UNROLL], y[UNROLL];
        uint32_t f[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nchunks) {
                x[u] = a4[c];
                y[u] = b4[c];
                f[u] = (LOGIC == SDB_NONE) ? 0u : in32[c];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nchunks) {
                const int32_t xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
                const int32_t ys[4] = {y[u].x, y[u].y, y[u].z, y[u].w};
                uint32_t r = 0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    r |= (uint32_t)logical(LOGIC, byte_of(f[u], k) != 0, compare(CMP, xs[k], ys[k])) << (8 * k);
                out32[c] = r;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
        const int64_t i = (nchunks << 2) + threadIdx.x;
        out[i] = logical(LOGIC, in[i] != 0, compare(CMP, a[i], b[i]));
    }
}

template <int CMP, int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_select_col_row(const int32_t *a, const int32_t *b,
                                                          const uint8_t *in, uint8_t *out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = logical(LOGIC, in[i] != 0, compare(CMP, a[i], b[i]));
}

template <int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_logical(const uint8_t *in, const uint8_t *local,
                                                   uint8_t *out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = logical(LOGIC, in[i] != 0, local[i] != 0);
}

// 0x01 in each byte of x that is nonzero, 0x00 elsewhere (no carry crosses a
// byte: (b & 0x7f) + 0x7f <= 0xfe)
__device__ __forceinline__ uint32_t nz_bytes(uint32_t x)
{
    return ((((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) >> 7) & 0x01010101u;
}

template <int LOGIC> __device__ __forceinline__ uint32_t logical_bytes(uint32_t a, uint32_t b)
{
    return LOGIC == SDB_NONE ? nz_bytes(b) : LOGIC == SDB_AND ? (nz_bytes(a) & nz_bytes(b)) : (nz_bytes(a) | nz_bytes(b));
}

// The same, 16 flags per thread and step (the three buffers 16-byte
// aligned); the last n % 16 flags one per thread.
template <int LOGIC>
__global__ void __launch_bounds__(BLOCK) k_logical_vec(const uint8_t *__restrict__ in,
                                                       const uint8_t *__restrict__ local, uint8_t *__restrict__ out,
                                                       int64_t n)
{
    const int64_t nq = n >> 4;
    const uint4 *a4 = reinterpret_cast<const uint4 *>(in);
    const uint4 *b4 = reinterpret_cast<const uint4 *>(local);
    uint4 *o4 = reinterpret_cast<uint4 *>(out);
    for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nq; base += (int64_t)gridDim.x * BLOCK * UNROLL) {
        uint4 a[UNROLL], b[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nq) {
                a[u] = LOGIC == SDB_NONE ? make_uint4(0, 0, 0, 0) : a4[c];
                b[u] = b4[c];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            if (c < nq)
                o4[c] = make_uint4(logical_bytes<LOGIC>(a[u].x, b[u].x), logical_bytes<LOGIC>(a[u].y, b[u].y),
                                   logical_bytes<LOGIC>(a[u].z, b[u].z), logical_bytes<LOGIC>(a[u].w, b[u].w));
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 15)) {
        const int64_t i = (nq << 4) + threadIdx.x;
        out[i] = logical(LOGIC, in[i] != 0, local[i] != 0);
    }
}

// ------------------------------------------------------------ projection

__global__ void __launch_bounds__(BLOCK) k_fill_i32(int32_t *dst, int32_t v, int64_t n)
{
    const int64_t head = ((16 - ((uintptr_t)dst & 15)) & 15) / 4; // rows before 16B boundary
    const int64_t h = head < n ? head : n;
    if (blockIdx.x == 0 && threadIdx.x < h) dst[threadIdx.x] = v;
    int32_t *body = dst + h;
    const int64_t m = n - h;
    const int64_t nchunks = m >> 2;
    int4 *d4 = reinterpret_cast<int4 *>(body);
    const int4 vv = make_int4(v, v, v, v);
    for (int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < nchunks; c += (int64_t)gridDim.x * BLOCK)
        d4[c] = vv;
    if (blockIdx.x == 0 && threadIdx.x < (m & 3)) body[(nchunks << 2) + threadIdx.x] = v;
}

// 16-byte stores between the first and the last 16-byte boundary of the
// buffer; the head and tail bytes (fewer than 16 each) one per thread
__global__ void __launch_bounds__(BLOCK) k_fill_u8(uint8_t *dst, uint8_t v, int64_t n)
{
    const int64_t head = min(n, (int64_t)((16u - ((uintptr_t)dst & 15u)) & 15u));
    const int64_t nq = (n - head) >> 4, tail0 = head + (nq << 4);
    const uint32_t w = 0x01010101u * v;
    const uint4 q = make_uint4(w, w, w, w);
    uint4 *dq = reinterpret_cast<uint4 *>(dst + head);
    const int64_t tid = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    for (int64_t i = tid; i < nq; i += (int64_t)gridDim.x * BLOCK) dq[i] = q;
    if (tid < head) dst[tid] = v;
    if (tid < n - tail0) dst[tail0 + tid] = v;
}

__global__ void k_zero_u64(uint64_t *p) { *p = 0; }

// kind: 0 = column op column, 1 = literal op column, 2 = column op literal
template <int OP, int KIND>
__global__ void __launch_bounds__(BLOCK) k_binop(int32_t *result, const int32_t *a, const int32_t *b,
                                                 int32_t lit, const uint8_t *flags, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        if (flags[i]) {
            const int32_t x = KIND == 1 ? lit : a[i];
            const int32_t y = KIND == 2 ? lit : b[i];
            result[i] = element_op(OP, x, y);
        }
    }
}

// Vector form: 4 rows per chunk, rows whose flag is 0 keep their content.
template <int OP, int KIND>
__global__ void __launch_bounds__(BLOCK) k_binop_vec(int32_t *result, const int32_t *a,
                                                     const int32_t *b, int32_t lit,
                                                     const uint8_t *flags, int64_t n)
{
    const int64_t nchunks = n >> 2;
    const uint32_t *f32 = reinterpret_cast<const uint32_t *>(flags);
    for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
         base += (int64_t)gridDim.x * BLOCK * UNROLL) {
        uint32_t f[UNROLL];
        int4 x[UNROLL], y[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t c = base + u * BLOCK + threadIdx.x;
            f[u] = c < nchunks ? f32[c] : 0u;
            if (f[u]) {
                if (KIND != 1) x[u] = reinterpret_cast<const int4 *>(a)[c];
                if (KIND != 2) y[u] = reinterpret_cast<const int4 *>(b)[c];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            if (!f[u]) continue;
            const int64_t c = base + u * BLOCK + threadIdx.x;
            const int32_t xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
            const int32_t ys[4] = {y[u].x, y[u].y, y[u].z, y[u].w};
            int32_t r[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                r[k] = element_op(OP, KIND == 1 ? lit : xs[k], KIND == 2 ? lit : ys[k]);
            if (f[u] == 0x01010101u) {
                reinterpret_cast<int4 *>(result)[c] = make_int4(r[0], r[1], r[2], r[3]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (byte_of(f[u], k)) result[4 * c + k] = r[k];
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
        const int64_t i = (nchunks << 2) + threadIdx.x;
        if (flags[i]) result[i] = element_op(OP, KIND == 1 ? lit : a[i], KIND == 2 ? lit : b[i]);
    }
}

// ------------------------------------------------------------------ join

__global__ void __launch_bounds__(BLOCK) k_build_keys(const int32_t *keys, const uint8_t *flags,
                                                      int64_t n, uint8_t *ht, int32_t len,
                                                      int32_t min)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const uint32_t s = slot_of(keys[i], min);
        if (s < (uint32_t)len) ht[s] = flags[i] ? 1 : 0;
    }
}

__global__ void __launch_bounds__(BLOCK) k_build_keyvals(const int32_t *keys, const int32_t *vals,
                                                         const uint8_t *flags, int64_t n,
                                                         int32_t *ht, int32_t len, int32_t min)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const uint32_t s = slot_of(keys[i], min);
        if (s >= (uint32_t)len) continue;
        if (flags[i])
            reinterpret_cast<int2 *>(ht)[s] = make_int2(1, vals[i]);
        else
            ht[2 * (int64_t)s] = 0;
    }
}

// vec: 4 rows per thread and step (keys 16-byte, flags 4-byte aligned),
// else one row.  The keys load beside the flags, not after them: one wait
// fewer per step, for the keys of chunks without a flagged row.
__global__ void __launch_bounds__(BLOCK) k_probe_semi(const int32_t *__restrict__ probe,
                                                      const uint8_t *in, uint8_t *out, int64_t n,
                                                      const uint8_t *__restrict__ ht, int32_t min,
                                                      int32_t max, int vec)
{
    auto present = [&](int32_t x) { return x >= min && x <= max && ht[slot_of(x, min)] != 0; };
    if (vec) {
        const int64_t nchunks = n >> 2;
        const int4 *p4 = reinterpret_cast<const int4 *>(probe);
        const uint32_t *in32 = reinterpret_cast<const uint32_t *>(in);
        uint32_t *out32 = reinterpret_cast<uint32_t *>(out);
        for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
             base += (int64_t)gridDim.x * BLOCK * UNROLL) {
            uint32_t f[UNROLL];
            int4 x[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * BLOCK + threadIdx.x;
                f[u] = c < nchunks ? in32[c] : 0u;
                x[u] = c < nchunks ? p4[c] : make_int4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * BLOCK + threadIdx.x;
                if (c >= nchunks) continue;
                uint32_t r = 0;
                if (f[u]) {
                    const int32_t xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const bool hit = byte_of(f[u], k) && present(xs[k]);
                        r |= (uint32_t)hit << (8 * k);
                    }
                }
                out32[c] = r;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
            const int64_t i = (nchunks << 2) + threadIdx.x;
            out[i] = in[i] && present(probe[i]);
        }
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const bool f = in[i] != 0;
        bool r = false;
        if (f) r = present(probe[i]);
        out[i] = r;
    }
}

// The value table of a probe: the reference's (flag, value) pairs, or --
// when the caller's table is much smaller than the probe column -- a 16-bit
// copy built on the device just before (k_pairs_range, k_pairs_to_u16): one
// uint16 per slot, value - vmin where the flag is 1, VALS_UNSET elsewhere.
// A quarter of the bytes per gather keeps a table of millions of slots in
// L2; the device decides (info[2]) whether every value fitted, so the host
// never waits for it.
struct PairsOrU16 {
    const int2 *pairs;
    const uint16_t *u16; // null: pairs only
    const int32_t *info; // vmin, vmax, values do not fit in 16 bits
};

__device__ __forceinline__ bool probe_value(const PairsOrU16 &t, bool use16, int32_t vmin, uint32_t slot, int32_t &v)
{
    if (use16) {
        const uint32_t x = t.u16[slot];
        v = vmin + (int32_t)x;
        return x != VALS_UNSET;
    }
    const int2 e = t.pairs[slot];
    v = e.y;
    return e.x == 1;
}

__global__ void __launch_bounds__(BLOCK) k_probe_full(const int32_t *__restrict__ probe,
                                                      int32_t *val_out, const uint8_t *in,
                                                      uint8_t *out, int64_t n,
                                                      PairsOrU16 t, int32_t min,
                                                      int32_t max, int vec)
{
    const bool use16 = t.u16 && t.info[2] == 0;
    const int32_t vmin = t.u16 ? t.info[0] : 0;
    if (vec) {
        const int64_t nchunks = n >> 2;
        const int4 *p4 = reinterpret_cast<const int4 *>(probe);
        const uint32_t *in32 = reinterpret_cast<const uint32_t *>(in);
        uint32_t *out32 = reinterpret_cast<uint32_t *>(out);
        for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
             base += (int64_t)gridDim.x * BLOCK * UNROLL) {
            uint32_t f[UNROLL];
            int4 x[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * BLOCK + threadIdx.x;
                f[u] = c < nchunks ? in32[c] : 0u;
                if (f[u]) x[u] = p4[c];
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * BLOCK + threadIdx.x;
                if (c >= nchunks) continue;
                uint32_t r = 0;
                if (f[u]) {
                    const int32_t xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        int32_t v;
                        if (byte_of(f[u], k) && xs[k] >= min && xs[k] <= max &&
                            probe_value(t, use16, vmin, slot_of(xs[k], min), v)) {
                            val_out[4 * c + k] = v;
                            r |= 1u << (8 * k);
                        }
                    }
                }
                out32[c] = r;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
            const int64_t i = (nchunks << 2) + threadIdx.x;
            bool r = false;
            if (in[i]) {
                const int32_t x = probe[i];
                int32_t v;
                if (x >= min && x <= max && probe_value(t, use16, vmin, slot_of(x, min), v)) {
                    val_out[i] = v;
                    r = true;
                }
            }
            out[i] = r;
        }
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        bool r = false;
        if (in[i]) {
            const int32_t x = probe[i];
            int32_t v;
            if (x >= min && x <= max && probe_value(t, use16, vmin, slot_of(x, min), v)) {
                val_out[i] = v;
                r = true;
            }
        }
        out[i] = r;
    }
}

__global__ void k_pairs_range_init(int32_t *info)
{
    info[0] = INT32_MAX;
    info[1] = INT32_MIN;
    info[2] = 0;
}

// min and max of the values of the slots whose flag is 1
__global__ void __launch_bounds__(BLOCK) k_pairs_range(const int2 *pairs, int64_t len, int32_t *info)
{
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t h = (int64_t)blockIdx.x * BLOCK + threadIdx.x; h < len; h += (int64_t)gridDim.x * BLOCK) {
        const int2 e = pairs[h];
        if (e.x == 1) {
            lo = min(lo, e.y);
            hi = max(hi, e.y);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    // one atomic pair per block: same-address atomics from every wave
    // serialise in L2
    __shared__ int32_t wlo[BLOCK / 64], whi[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        wlo[threadIdx.x >> 6] = lo;
        whi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; w++) {
            lo = min(lo, wlo[w]);
            hi = max(hi, whi[w]);
        }
        if (lo <= hi) {
            atomicMin(&info[0], lo);
            atomicMax(&info[1], hi);
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_pairs_to_u16(const int2 *pairs, int64_t len, int32_t *info, uint16_t *u16)
{
    const int32_t vmin = info[0];
    // every value offset must stay below VALS_UNSET
    const bool fits = info[1] < info[0] || (int64_t)info[1] - (int64_t)vmin < (int64_t)VALS_UNSET;
    if (!fits) {
        if (blockIdx.x == 0 && threadIdx.x == 0) info[2] = 1;
        return;
    }
    for (int64_t h = (int64_t)blockIdx.x * BLOCK + threadIdx.x; h < len; h += (int64_t)gridDim.x * BLOCK) {
        const int2 e = pairs[h];
        u16[h] = e.x == 1 ? (uint16_t)(e.y - vmin) : (uint16_t)VALS_UNSET;
    }
}

// ----------------------------------------------------------- aggregation

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum; one atomic per block into *acc.
__device__ __forceinline__ void block_add_u64(uint64_t v, uint64_t *acc)
{
    __shared__ uint64_t part[BLOCK / 64];
    v = wave_sum_u64(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; w++) s += part[w];
        if (s) atomicAdd(reinterpret_cast<unsigned long long *>(acc), (unsigned long long)s);
    }
}

__global__ void __launch_bounds__(BLOCK) k_sum_flagged(const int32_t *__restrict__ col,
                                                       const uint8_t *__restrict__ flags, int64_t n,
                                                       uint64_t *acc, int vec)
{
    int64_t s = 0; // int64 partial; the uint64 total is its two's-complement image
    if (vec) {
        const int64_t nchunks = n >> 2;
        const int4 *c4 = reinterpret_cast<const int4 *>(col);
        const uint32_t *f32 = reinterpret_cast<const uint32_t *>(flags);
        for (int64_t base = (int64_t)blockIdx.x * BLOCK * UNROLL; base < nchunks;
             base += (int64_t)gridDim.x * BLOCK * UNROLL) {
            uint32_t f[UNROLL];
            int4 x[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * BLOCK + threadIdx.x;
                f[u] = c < nchunks ? f32[c] : 0u;
                if (f[u]) x[u] = c4[c];
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                if (!f[u]) continue;
                s += byte_of(f[u], 0) ? (int64_t)x[u].x : 0;
                s += byte_of(f[u], 1) ? (int64_t)x[u].y : 0;
                s += byte_of(f[u], 2) ? (int64_t)x[u].z : 0;
                s += byte_of(f[u], 3) ? (int64_t)x[u].w : 0;
            }
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
            const int64_t i = (nchunks << 2) + threadIdx.x;
            if (flags[i]) s += col[i];
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
            if (flags[i]) s += col[i];
    }
    block_add_u64((uint64_t)s, acc);
}

__global__ void __launch_bounds__(BLOCK) k_count_flags(const uint8_t *flags, int64_t n, uint64_t *acc,
                                                       int vec)
{
    uint64_t s = 0;
    if (vec) {
        const int64_t nw = n >> 2;
        const uint32_t *f32 = reinterpret_cast<const uint32_t *>(flags);
        for (int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < nw; c += (int64_t)gridDim.x * BLOCK) {
            const uint32_t w = f32[c];
            // one bit per non-zero byte
            const uint32_t nz = (((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w) & 0x80808080u;
            s += __popc(nz);
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) s += flags[(nw << 2) + threadIdx.x] != 0;
    } else {
        for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
            s += flags[i] != 0;
    }
    block_add_u64(s, acc);
}

constexpr int GB_MAX_KEYS = 8;
struct GroupArgs {
    const int32_t *keys[GB_MAX_KEYS];
    int32_t *out_keys[GB_MAX_KEYS];
    int32_t mins[GB_MAX_KEYS];
    int32_t maxs[GB_MAX_KEYS];
    uint32_t range[GB_MAX_KEYS]; // max - min + 1 (uint32 image of the int)
    int k;
};

// Dense slot of row i exactly as GroupByAggregateKernel computes it
// (int32 wrapping accumulation, then % prod on the uint64 image).
__device__ __forceinline__ uint32_t group_slot(const GroupArgs &g, const int32_t *kv, uint64_t prod)
{
    uint32_t hash = 0, mult = 1;
    for (int j = 0; j < g.k; j++) {
        hash += ((uint32_t)kv[j] - (uint32_t)g.mins[j]) * mult;
        mult *= g.range[j];
    }
    const uint64_t h64 = (uint64_t)(int64_t)(int32_t)hash;
    return (uint32_t)(h64 % prod);
}

// Global path: one exchange + one atomic add per flagged row.
__device__ __forceinline__ void group_row_global(const GroupArgs &g, const int32_t *kv, int32_t agg,
                                                 uint64_t *sums, uint32_t *seen, uint64_t prod)
{
    const uint32_t s = group_slot(g, kv, prod);
    if (atomicExch(&seen[s], 1u) == 0u)
        for (int j = 0; j < g.k; j++) g.out_keys[j][s] = kv[j];
    atomicAdd(reinterpret_cast<unsigned long long *>(&sums[s]), (unsigned long long)(int64_t)agg);
}

__global__ void __launch_bounds__(BLOCK) k_groupby_global(GroupArgs g, const int32_t *agg,
                                                          const uint8_t *flags, int64_t n,
                                                          uint64_t *sums, uint32_t *seen, uint64_t prod)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        if (!flags[i]) continue;
        int32_t kv[GB_MAX_KEYS];
        for (int j = 0; j < g.k; j++) kv[j] = g.keys[j][i];
        group_row_global(g, kv, agg[i], sums, seen, prod);
    }
}

// LDS-privatised dense group-by: each block accumulates its rows into an LDS
// copy of the dense sum array (ds_add_u64) plus an occupancy bitmap, then
// flushes the occupied slots with one global atomic each.  Rows whose keys
// all lie in [min, max] map to a unique slot whose key values are recovered
// from the slot index at flush; any other row takes the global path so the
// key it stores is its own.
constexpr int GB_BLOCK = 1024; // threads per block of the LDS group-by

__global__ void __launch_bounds__(GB_BLOCK) k_groupby_lds(GroupArgs g, const int32_t *agg,
                                                          const uint8_t *flags, int64_t n,
                                                          uint64_t *sums, uint32_t *seen,
                                                          uint32_t prod, int vec)
{
    extern __shared__ unsigned long long lds[];
    unsigned long long *lsum = lds;
    uint32_t *lbits = reinterpret_cast<uint32_t *>(lds + prod);
    const uint32_t nwords = (prod + 31) / 32;
    for (uint32_t s = threadIdx.x; s < prod; s += GB_BLOCK) lsum[s] = 0;
    for (uint32_t w = threadIdx.x; w < nwords; w += GB_BLOCK) lbits[w] = 0;
    __syncthreads();
    auto row = [&](int64_t i) {
        int32_t kv[GB_MAX_KEYS];
        bool in_range = true;
        for (int j = 0; j < g.k; j++) {
            kv[j] = g.keys[j][i];
            in_range &= kv[j] >= g.mins[j] && kv[j] <= g.maxs[j];
        }
        if (!in_range) {
            group_row_global(g, kv, agg[i], sums, seen, prod);
            return;
        }
        const uint32_t s = group_slot(g, kv, prod);
        atomicAdd(&lsum[s], (unsigned long long)(int64_t)agg[i]);
        atomicOr(&lbits[s >> 5], 1u << (s & 31));
    };
    if (vec) {
        // four flags per load (4-byte aligned), UNROLL loads in flight per
        // thread; only flagged rows read their keys and value
        const int64_t nchunks = n >> 2;
        const uint32_t *f32 = reinterpret_cast<const uint32_t *>(flags);
        for (int64_t base = (int64_t)blockIdx.x * GB_BLOCK * UNROLL; base < nchunks;
             base += (int64_t)gridDim.x * GB_BLOCK * UNROLL) {
            uint32_t f[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int64_t c = base + u * GB_BLOCK + threadIdx.x;
                f[u] = c < nchunks ? f32[c] : 0u;
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                if (!f[u]) continue;
                const int64_t c = base + u * GB_BLOCK + threadIdx.x;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (byte_of(f[u], k)) row(4 * c + k);
            }
        }
        if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
            const int64_t i = (nchunks << 2) + threadIdx.x;
            if (flags[i]) row(i);
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * GB_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * GB_BLOCK)
            if (flags[i]) row(i);
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < prod; s += GB_BLOCK) {
        if (!((lbits[s >> 5] >> (s & 31)) & 1u)) continue;
        atomicAdd(reinterpret_cast<unsigned long long *>(&sums[s]), lsum[s]);
        if (atomicExch(&seen[s], 1u) == 0u) {
            uint32_t rest = s;
            for (int j = 0; j < g.k; j++) {
                g.out_keys[j][s] = (int32_t)((uint32_t)g.mins[j] + rest % g.range[j]);
                rest /= g.range[j];
            }
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < (int64_t)prod; i += (int64_t)gridDim.x * BLOCK)
        flags[i] = seen[i] != 0;
}

// ------------------------------------------------------ compact probe tables
// Deterministic (no atomics on the tables).  The population count is left
// per block and summed when the statistics are read (k_stats_reduce).

// Stores the block's count (the sum of every wave's, held by its lane 0) in
// partials[blockIdx.x].  One same-address atomic per block instead
// serialises: 1,368 blocks over the 1.4M-slot part table took 21 us, 7 us
// with the partial counts.
__device__ __forceinline__ void block_count_store(uint32_t local, uint32_t *partials)
{
    __shared__ uint32_t wave_sum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_sum[wave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; w++) t += wave_sum[w];
        partials[blockIdx.x] = t;
    }
}

// Presence bitmaps of the reference tables: one thread per slot, each wave
// covering 64 consecutive slots (coalesced reads), its ballot giving two
// bitmap words; the set slots are counted per wave, then per block (one
// partial count per block, bits_partials(len) of them).
// Each wave takes BITS_GROUPS consecutive 64-slot groups per step, their
// loads issued together, on a grid of at most BITS_MAX_BLOCKS blocks.  From
// the pairs, the 16-bit value table is written too: value - vmin in the set
// slots, VALS_UNSET in the others (`bad` is raised when a set slot's offset
// does not fit below VALS_UNSET, and the table is then not used).
constexpr int BITS_GROUPS = 4;
constexpr unsigned BITS_MAX_BLOCKS = 2048;

__global__ void __launch_bounds__(BLOCK) k_bits_from_keys(const uint8_t *ht, int32_t len, uint32_t *bits,
                                                          uint32_t *partials)
{
    const int64_t words = ((int64_t)len + 31) / 32;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
    uint32_t local = 0;
    for (int64_t g0 = (((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6) * BITS_GROUPS; g0 * 64 < (int64_t)len;
         g0 += nwaves * BITS_GROUPS) {
        bool set[BITS_GROUPS];
#pragma unroll
        for (int k = 0; k < BITS_GROUPS; k++) {
            const int64_t s = (g0 + k) * 64 + lane;
            set[k] = s < len && ht[s] != 0;
        }
#pragma unroll
        for (int k = 0; k < BITS_GROUPS; k++) {
            const int64_t g = g0 + k;
            const uint64_t m = __ballot(set[k]);
            if (lane == 0 && 2 * g < words) {
                bits[2 * g] = (uint32_t)m;
                if (2 * g + 1 < words) bits[2 * g + 1] = (uint32_t)(m >> 32);
                local += (uint32_t)__popcll(m);
            }
        }
    }
    block_count_store(local, partials);
}

__global__ void __launch_bounds__(BLOCK) k_bits_from_pairs(const int2 *ht, int32_t len, uint32_t *bits,
                                                           uint16_t *vals, int32_t vmin, uint32_t *partials,
                                                           uint32_t *bad)
{
    const int64_t words = ((int64_t)len + 31) / 32;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * BLOCK) >> 6;
    uint32_t local = 0;
    bool ok = true;
    for (int64_t g0 = (((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6) * BITS_GROUPS; g0 * 64 < (int64_t)len;
         g0 += nwaves * BITS_GROUPS) {
        int2 e[BITS_GROUPS];
#pragma unroll
        for (int k = 0; k < BITS_GROUPS; k++) {
            const int64_t s = (g0 + k) * 64 + lane;
            e[k] = s < len ? ht[s] : make_int2(0, 0);
        }
#pragma unroll
        for (int k = 0; k < BITS_GROUPS; k++) {
            const int64_t g = g0 + k, s = g * 64 + lane;
            const bool set = e[k].x == 1;
            // unset slots hold VALS_UNSET; a set slot's offset must stay below it
            const uint32_t d = set ? (uint32_t)e[k].y - (uint32_t)vmin : VALS_UNSET;
            ok &= !set || d < VALS_UNSET;
            if (vals && s < len) vals[s] = (uint16_t)d;
            const uint64_t m = __ballot(set);
            if (lane == 0 && 2 * g < words) {
                bits[2 * g] = (uint32_t)m;
                if (2 * g + 1 < words) bits[2 * g + 1] = (uint32_t)(m >> 32);
                local += (uint32_t)__popcll(m);
            }
        }
    }
    block_count_store(local, partials);
    if (!ok) atomicOr(bad, 1u);
}

// The windows of 32 consecutive keys (one bitmap word) of a table of WB-byte
// windows (1, 2 or 4 bytes), 16-byte aligned and padded to 16 bytes, in
// 2 * WB vector loads issued together.
template <uint32_t WB> struct DictWord {
    static constexpr uint32_t NQ = 2 * WB; // 16-byte chunks per 32 windows
    static constexpr uint32_t FIELD = WB >= 4 ? ~0u : (1u << (8u * (WB % 4u))) - 1u;
    uint4 q[NQ];
    // the windows of keys [32 w, 32 w + 32) that are below len; a chunk that
    // starts past the table's last window is not read
    __device__ __forceinline__ void load(const uint8_t *tab, uint32_t len, uint32_t w)
    {
        const uint4 *p = reinterpret_cast<const uint4 *>(tab) + (size_t)w * NQ;
        const uint32_t nbytes = min(len - w * 32u, 32u) * WB;
#pragma unroll
        for (uint32_t k = 0; k < NQ; k++) q[k] = 16u * k < nbytes ? p[k] : make_uint4(0u, 0u, 0u, 0u);
    }
    __device__ __forceinline__ uint32_t window(uint32_t e) const // e: compile-time after unrolling
    {
        const uint32_t byte = e * WB, k = byte >> 4, j = (byte >> 2) & 3u;
        const uint32_t c = j == 0 ? q[k].x : j == 1 ? q[k].y : j == 2 ? q[k].z : q[k].w;
        return (c >> (8u * (byte & 3u))) & FIELD;
    }
    __device__ __forceinline__ uint32_t bits(uint32_t len, uint32_t w) const
    {
        uint32_t m = 0;
#pragma unroll
        for (uint32_t e = 0; e < 32u; e++) m |= (uint32_t)(window(e) != 0u) << e;
        const uint32_t n = min(len - w * 32u, 32u);
        return n == 32u ? m : m & ((1u << n) - 1u);
    }
    // the windows of the keys set in m, at ranks rank, rank + 1, ... (those
    // at or past cap are dropped)
    template <class T> __device__ __forceinline__ void write(uint32_t m, uint32_t rank, T *vals, uint32_t cap) const
    {
#pragma unroll
        for (uint32_t e = 0; e < 32u; e++)
            if ((m >> e) & 1u) {
                if (rank < cap) vals[rank] = (T)window(e);
                rank++;
            }
    }
};

// Rank dictionary of a presence bitmap with per-slot 16-bit values (the
// layout FV_RANK16 describes in fused_types.hpp), from the per-block counts
// of k_bits_from_pairs: its block b counted exactly the slots
// [1024 b, 1024 b + 1024), i.e. the bitmap words [32 b, 32 b + 32).  One wave
// per such range: its set slots before the range are the partial counts of
// the blocks before it; lane l < 32 takes word 32 b + l, the wave scans the
// words' counts, and each lane writes its word, its prefix and its set
// slots' values (four vector loads issued together).
constexpr uint32_t RANK_PART_WORDS = 32;

__global__ void __launch_bounds__(64) k_rank_dict_parts(const uint32_t *__restrict__ bits, int32_t words,
                                                        const uint32_t *__restrict__ partials,
                                                        const uint16_t *__restrict__ vals, uint32_t len,
                                                        uint8_t *__restrict__ dict, uint32_t cap_vals)
{
    const int lane = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const int32_t w = (int32_t)(b * RANK_PART_WORDS) + (lane & 31);
    const uint32_t m = lane < 32 && w < words ? bits[w] : 0u;
    uint32_t before = 0;
    for (uint32_t j = lane; j < b; j += 64) before += partials[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) before += __shfl_xor(before, off, 64);
    const uint32_t c = __popc(m);
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane >= 32 || w >= words) return;
    const uint32_t base = before + incl - c;
    reinterpret_cast<uint32_t *>(dict)[w] = m;
    reinterpret_cast<uint16_t *>(dict + rank_prefix_off((uint32_t)words))[w] = (uint16_t)base;
    if (m) {
        DictWord<2> x;
        x.load(reinterpret_cast<const uint8_t *>(vals), len, (uint32_t)w);
        x.write(m, base, reinterpret_cast<uint16_t *>(dict + rank_values_off((uint32_t)words)), cap_vals);
    }
}

// ------------------------------------------------------- batch probe masks
// One thread per key of a masked fact column (blockIdx.y = column): the key's
// word starts with every program's bit and loses the bit of each program one
// of whose probes on the column fails for the key -- the probe's own test
// (range check, then the table entry: FilterJoinKernel / FullJoinKernel,
// kernels/join.hpp:129-160, 256-300).
__global__ void __launch_bounds__(BLOCK) k_batch_masks(const FMaskCol *cols, const FMaskEntry *ents)
{
    const FMaskCol c = cols[blockIdx.y];
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < c.len; i += gridDim.x * BLOCK) {
        const int32_t key = (int32_t)((uint32_t)c.lo + i);
        uint32_t w = c.full;
        for (uint32_t e = c.first; e < c.first + c.count; e++) {
            const FMaskEntry E = ents[e];
            bool pass = key >= E.ht_min && key <= E.ht_max;
            if (pass) {
                const uint32_t s = slot_of(key, E.ht_min);
                if (E.kind == FM_BITS) pass = (reinterpret_cast<const uint32_t *>(E.ht)[s >> 5] >> (s & 31)) & 1u;
                else if (E.kind == FM_SEMI) pass = reinterpret_cast<const uint8_t *>(E.ht)[s] != 0;
                else pass = reinterpret_cast<const int2 *>(E.ht)[s].x == 1;
            }
            if (!pass) w &= ~(1u << E.bit);
        }
        const uint32_t v = w >> c.shift;
        if (c.wbytes == 1) c.tab[i] = (uint8_t)v;
        else if (c.wbytes == 2) reinterpret_cast<uint16_t *>(c.tab)[i] = (uint16_t)v;
        else reinterpret_cast<uint32_t *>(c.tab)[i] = v;
    }
}

// Rank dictionaries of batched mask tables (FMaskDict), one block per table.
// Each thread takes whole bitmap words: the word's 32 windows (32 x 1, 2 or 4
// bytes, 16-byte aligned -- every table starts on 16 bytes and is padded to
// 16) come in with 2, 4 or 8 vector loads issued together, and the word's
// bit per key is whether the key's window is nonzero.  The exclusive counts
// before each word are scanned across the block; then each thread reloads its
// words' windows and writes every set key's window at its rank.
constexpr int DICT_BLOCK = 256;
constexpr uint32_t DICT_MAX_WORDS = F_DICT_MAX_LEN / 32;

template <uint32_t WB> __device__ void dict_bits(const FMaskDict &d, uint32_t words, uint32_t *sbits)
{
    for (uint32_t w = threadIdx.x; w < words; w += DICT_BLOCK) {
        DictWord<WB> x;
        x.load(d.tab, d.len, w);
        sbits[w] = x.bits(d.len, w);
    }
}

template <uint32_t WB>
__device__ void dict_values(const FMaskDict &d, uint32_t words, const uint32_t *sbits, const uint32_t *spre,
                            uint32_t *vals)
{
    for (uint32_t w = threadIdx.x; w < words; w += DICT_BLOCK) {
        const uint32_t m = sbits[w];
        if (!m) continue;
        DictWord<WB> x;
        x.load(d.tab, d.len, w);
        x.write(m, spre[w], vals, ~0u);
    }
}

__global__ void __launch_bounds__(DICT_BLOCK) k_batch_dict(const FMaskDict *dicts)
{
    __shared__ uint32_t sbits[DICT_MAX_WORDS];
    __shared__ uint32_t spre[DICT_MAX_WORDS];
    __shared__ uint32_t wave_total[DICT_BLOCK / 64];
    const FMaskDict d = dicts[blockIdx.x];
    const uint32_t words = dict_words(d.len);
    if (d.wbytes == 1) dict_bits<1>(d, words, sbits);
    else if (d.wbytes == 2) dict_bits<2>(d, words, sbits);
    else dict_bits<4>(d, words, sbits);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t per = (words + DICT_BLOCK - 1) / DICT_BLOCK;
    const uint32_t a = min(words, threadIdx.x * per), b = min(words, a + per);
    uint32_t local = 0;
    for (uint32_t w = a; w < b; w++) local += __popc(sbits[w]);
    uint32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    uint32_t base = incl - local, total = 0;
    for (int v = 0; v < DICT_BLOCK / 64; v++) {
        if (v < wave) base += wave_total[v];
        total += wave_total[v];
    }
    for (uint32_t w = a; w < b; w++) {
        spre[w] = base;
        base += __popc(sbits[w]);
    }
    __syncthreads();
    uint32_t *hdr = reinterpret_cast<uint32_t *>(d.dict);
    uint32_t *obits = reinterpret_cast<uint32_t *>(d.dict + 16);
    uint16_t *opre = reinterpret_cast<uint16_t *>(d.dict + 16 + dict_prefix_off(d.len));
    uint32_t *ovals = reinterpret_cast<uint32_t *>(d.dict + 16 + dict_vals_off(d.len));
    for (uint32_t w = threadIdx.x; w < words; w += DICT_BLOCK) {
        obits[w] = sbits[w];
        opre[w] = (uint16_t)spre[w];
    }
    const bool ok = total <= F_DICT_CAP;
    if (ok) {
        if (d.wbytes == 1) dict_values<1>(d, words, sbits, spre, ovals);
        else if (d.wbytes == 2) dict_values<2>(d, words, sbits, spre, ovals);
        else dict_values<4>(d, words, sbits, spre, ovals);
    }
    if (threadIdx.x == 0) {
        hdr[0] = ok ? 1u : 0u;
        hdr[1] = total;
    }
}

// ------------------------------------------------------------ compaction
// Three passes over tiles of COMPACT_TILE rows: count per tile, exclusive
// scan of the counts (one block), then write ids in row order.
constexpr int COMPACT_PER_THREAD = 16;
constexpr int64_t COMPACT_TILE = (int64_t)BLOCK * COMPACT_PER_THREAD;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ void __launch_bounds__(BLOCK) k_compact_count(const uint8_t *flags, int64_t n, uint32_t *counts)
{
    const int64_t t0 = (int64_t)blockIdx.x * COMPACT_TILE + (int64_t)threadIdx.x * COMPACT_PER_THREAD;
    uint32_t c = 0;
    for (int k = 0; k < COMPACT_PER_THREAD; k++)
        if (t0 + k < n) c += flags[t0 + k] != 0;
    uint32_t total;
    block_exclusive_scan(c, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

__global__ void __launch_bounds__(BLOCK) k_compact_scan(uint32_t *counts, int64_t ntiles, uint32_t *count_out)
{
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < ntiles; base += BLOCK) {
        const int64_t t = base + threadIdx.x;
        const uint32_t v = t < ntiles ? counts[t] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, &total);
        if (t < ntiles) counts[t] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count_out = carry;
}

__global__ void __launch_bounds__(BLOCK) k_compact_write(const uint8_t *flags, int64_t n,
                                                         const uint32_t *offsets, int32_t *row_ids)
{
    const int64_t t0 = (int64_t)blockIdx.x * COMPACT_TILE + (int64_t)threadIdx.x * COMPACT_PER_THREAD;
    uint32_t c = 0;
    for (int k = 0; k < COMPACT_PER_THREAD; k++)
        if (t0 + k < n) c += flags[t0 + k] != 0;
    uint32_t total;
    uint32_t pos = offsets[blockIdx.x] + block_exclusive_scan(c, &total);
    for (int k = 0; k < COMPACT_PER_THREAD; k++)
        if (t0 + k < n && flags[t0 + k]) row_ids[pos++] = (int32_t)(t0 + k);
}

__global__ void __launch_bounds__(BLOCK) k_gather_cols(GatherCols g, const int32_t *ids, int64_t count)
{
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < count; j += (int64_t)gridDim.x * BLOCK) {
        const uint64_t r = (uint64_t)(uint32_t)ids[j];
        for (int c = 0; c < g.ncols; c++) {
            if ((g.is_u64 >> c) & 1u)
                static_cast<uint64_t *>(g.dst[c])[j] =
                    r < g.src_rows[c] ? static_cast<const uint64_t *>(g.src[c])[r] : 0ull;
            else
                static_cast<int32_t *>(g.dst[c])[j] = r < g.src_rows[c] ? static_cast<const int32_t *>(g.src[c])[r] : 0;
        }
    }
}

// blockIdx.y = part; the blocks of a part stride over its bytes (16 at a
// time where source and destination allow, the rest one by one)
__global__ void __launch_bounds__(BLOCK) k_pack_parts(PackParts p, uint8_t *dst)
{
    const int i = blockIdx.y;
    const uint8_t *src = p.src[i];
    uint8_t *out = dst + p.off[i];
    const uint64_t bytes = p.bytes[i];
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    uint64_t head = 0;
    if ((((uintptr_t)src | (uintptr_t)out) & 15) == 0) {
        head = bytes & ~uint64_t(15);
        for (uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < head / 16; j += stride)
            reinterpret_cast<uint4 *>(out)[j] = reinterpret_cast<const uint4 *>(src)[j];
    }
    for (uint64_t j = head + (uint64_t)blockIdx.x * BLOCK + threadIdx.x; j < bytes; j += stride) out[j] = src[j];
}

// The statistics of compact probe tables (one block per table): the sum of
// the table's partial counts and its overflow flag, as the pair dst[2 y],
// dst[2 y + 1].
__global__ void __launch_bounds__(BLOCK) k_stats_reduce(StatsParts p, uint32_t *dst)
{
    __shared__ uint32_t wave_sum[BLOCK / 64];
    const int y = blockIdx.y;
    const uint32_t *st = p.stats[y];
    uint32_t t = 0;
    for (uint32_t i = threadIdx.x; i < p.nparts[y]; i += BLOCK) t += st[1 + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_sum[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; w++) sum += wave_sum[w];
        dst[2 * y] = sum;
        dst[2 * y + 1] = st[0];
    }
}

__global__ void __launch_bounds__(BLOCK) k_gather_i32(const int32_t *src, const int32_t *ids, int64_t count,
                                                      int32_t *dst)
{
    for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < count; j += (int64_t)gridDim.x * BLOCK)
        dst[j] = src[ids[j]];
}

__global__ void __launch_bounds__(BLOCK) k_minmax(const int32_t *col, int64_t n, int32_t *mm)
{
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
        const int32_t v = col[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int32_t a = __shfl_xor(lo, off, 64), b = __shfl_xor(hi, off, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void k_minmax_init(int32_t *mm)
{
    mm[0] = INT32_MAX;
    mm[1] = INT32_MIN;
}

// --------------------------------------------------------- dispatch helpers

void check_launch(const char *what)
{
    hip_check(hipGetLastError(), what);
}

unsigned grid_rows(int64_t n) { return stream_grid(n, ROWS_PER_STEP); }
unsigned grid_one(int64_t n) { return stream_grid(n, BLOCK); }

template <int CMP, int LOGIC>
void run_select_lit(const int32_t *col, const uint8_t *in, uint8_t *out, int64_t n, int32_t v, hipStream_t s)
{
    if (aligned(col, 16) && aligned(in, 4) && aligned(out, 4))
        hipLaunchKernelGGL((k_select_lit<CMP, LOGIC>), dim3(grid_rows(n)), dim3(BLOCK), 0, s, col, in, out, n, v);
    else
        hipLaunchKernelGGL((k_select_lit_row<CMP, LOGIC>), dim3(grid_one(n)), dim3(BLOCK), 0, s, col, in, out, n, v);
}

template <int CMP, int LOGIC>
void run_select_col(const int32_t *a, const int32_t *b, const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s)
{
    if (aligned(a, 16) && aligned(b, 16) && aligned(in, 4) && aligned(out, 4))
        hipLaunchKernelGGL((k_select_col<CMP, LOGIC>), dim3(grid_rows(n)), dim3(BLOCK), 0, s, a, b, in, out, n);
    else
        hipLaunchKernelGGL((k_select_col_row<CMP, LOGIC>), dim3(grid_one(n)), dim3(BLOCK), 0, s, a, b, in, out, n);
}

template <int CMP>
void sel_lit_logic(int logic, const int32_t *col, const uint8_t *in, uint8_t *out, int64_t n, int32_t v, hipStream_t s)
{
    switch (logic) {
    case SDB_NONE: run_select_lit<CMP, SDB_NONE>(col, in, out, n, v, s); break;
    case SDB_AND: run_select_lit<CMP, SDB_AND>(col, in, out, n, v, s); break;
    default: run_select_lit<CMP, SDB_OR>(col, in, out, n, v, s); break;
    }
}

template <int CMP>
void sel_col_logic(int logic, const int32_t *a, const int32_t *b, const uint8_t *in, uint8_t *out, int64_t n, hipStream_t s)
{
    switch (logic) {
    case SDB_NONE: run_select_col<CMP, SDB_NONE>(a, b, in, out, n, s); break;
    case SDB_AND: run_select_col<CMP, SDB_AND>(a, b, in, out, n, s); break;
    default: run_select_col<CMP, SDB_OR>(a, b, in, out, n, s); break;
    }
}

template <int KIND>
void run_binop(int op, int32_t *r, const int32_t *a, const int32_t *b, int32_t lit, const uint8_t *f, int64_t n,
               hipStream_t s)
{
    const bool vec = aligned(r, 16) && aligned(f, 4) && (KIND == 1 || aligned(a, 16)) && (KIND == 2 || aligned(b, 16));
#define SDB_BINOP_CASE(OPV)                                                                              \
    case OPV:                                                                                            \
        if (vec)                                                                                         \
            hipLaunchKernelGGL((k_binop_vec<OPV, KIND>), dim3(grid_rows(n)), dim3(BLOCK), 0, s, r, a, b, lit, f, n); \
        else                                                                                             \
            hipLaunchKernelGGL((k_binop<OPV, KIND>), dim3(grid_one(n)), dim3(BLOCK), 0, s, r, a, b, lit, f, n); \
        break;
    switch (op) {
        SDB_BINOP_CASE(SDB_MUL)
        SDB_BINOP_CASE(SDB_DIV)
        SDB_BINOP_CASE(SDB_ADD)
        SDB_BINOP_CASE(SDB_SUB)
    }
#undef SDB_BINOP_CASE
}

} // namespace

// ------------------------------------------------------ internal launchers
// (used by the engine's unfused bundle path as well as the C ABI)

void launch_select_lit(const int32_t *col, const uint8_t *in, uint8_t *out, int64_t n, int cmp, int logic,
                       int32_t v, hipStream_t s)
{
    if (n <= 0) return;
    switch (cmp) {
    case SDB_EQ: sel_lit_logic<SDB_EQ>(logic, col, in, out, n, v, s); break;
    case SDB_NE: sel_lit_logic<SDB_NE>(logic, col, in, out, n, v, s); break;
    case SDB_LT: sel_lit_logic<SDB_LT>(logic, col, in, out, n, v, s); break;
    case SDB_LE: sel_lit_logic<SDB_LE>(logic, col, in, out, n, v, s); break;
    case SDB_GT: sel_lit_logic<SDB_GT>(logic, col, in, out, n, v, s); break;
    default: sel_lit_logic<SDB_GE>(logic, col, in, out, n, v, s); break;
    }
    check_launch("select_lit");
}

void launch_select_col(const int32_t *a, const int32_t *b, const uint8_t *in, uint8_t *out, int64_t n, int cmp,
                       int logic, hipStream_t s)
{
    if (n <= 0) return;
    switch (cmp) {
    case SDB_EQ: sel_col_logic<SDB_EQ>(logic, a, b, in, out, n, s); break;
    case SDB_NE: sel_col_logic<SDB_NE>(logic, a, b, in, out, n, s); break;
    case SDB_LT: sel_col_logic<SDB_LT>(logic, a, b, in, out, n, s); break;
    case SDB_LE: sel_col_logic<SDB_LE>(logic, a, b, in, out, n, s); break;
    case SDB_GT: sel_col_logic<SDB_GT>(logic, a, b, in, out, n, s); break;
    default: sel_col_logic<SDB_GE>(logic, a, b, in, out, n, s); break;
    }
    check_launch("select_col");
}

void launch_logical(const uint8_t *in, const uint8_t *local, uint8_t *out, int64_t n, int logic, hipStream_t s)
{
    if (n <= 0) return;
    if (aligned(in, 16) && aligned(local, 16) && aligned(out, 16)) {
        const unsigned grid = stream_grid((n >> 4) + 1, (int64_t)BLOCK * UNROLL);
        switch (logic) {
        case SDB_NONE: hipLaunchKernelGGL((k_logical_vec<SDB_NONE>), dim3(grid), dim3(BLOCK), 0, s, in, local, out, n); break;
        case SDB_AND: hipLaunchKernelGGL((k_logical_vec<SDB_AND>), dim3(grid), dim3(BLOCK), 0, s, in, local, out, n); break;
        default: hipLaunchKernelGGL((k_logical_vec<SDB_OR>), dim3(grid), dim3(BLOCK), 0, s, in, local, out, n); break;
        }
        check_launch("logical");
        return;
    }
    switch (logic) {
    case SDB_NONE: hipLaunchKernelGGL((k_logical<SDB_NONE>), dim3(grid_one(n)), dim3(BLOCK), 0, s, in, local, out, n); break;
    case SDB_AND: hipLaunchKernelGGL((k_logical<SDB_AND>), dim3(grid_one(n)), dim3(BLOCK), 0, s, in, local, out, n); break;
    default: hipLaunchKernelGGL((k_logical<SDB_OR>), dim3(grid_one(n)), dim3(BLOCK), 0, s, in, local, out, n); break;
    }
    check_launch("logical");
}

void launch_fill_i32(int32_t *dst, int32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(stream_grid(n, BLOCK * 4)), dim3(BLOCK), 0, s, dst, v, n);
    check_launch("fill_i32");
}

void launch_fill_u8(uint8_t *dst, uint8_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_one(std::max<int64_t>(n / 16, 1))), dim3(BLOCK), 0, s, dst, v, n);
    check_launch("fill_u8");
}

void launch_binop(int kind, int op, int32_t *r, const int32_t *a, const int32_t *b, int32_t lit,
                  const uint8_t *f, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    if (kind == 0) run_binop<0>(op, r, a, b, lit, f, n, s);
    else if (kind == 1) run_binop<1>(op, r, a, b, lit, f, n, s);
    else run_binop<2>(op, r, a, b, lit, f, n, s);
    check_launch("binop");
}

void launch_build_keys(const int32_t *keys, const uint8_t *flags, int64_t n, uint8_t *ht, int32_t len,
                       int32_t min, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_build_keys, dim3(grid_one(n)), dim3(BLOCK), 0, s, keys, flags, n, ht, len, min);
    check_launch("build_keys");
}

void launch_build_keyvals(const int32_t *keys, const int32_t *vals, const uint8_t *flags, int64_t n, int32_t *ht,
                          int32_t len, int32_t min, hipStream_t s)
{
    if (n <= 0) return;
    if (!aligned(ht, 8)) fail(SDB_ERR_INVALID_ARG, "key-value hash table must be 8-byte aligned");
    hipLaunchKernelGGL(k_build_keyvals, dim3(grid_one(n)), dim3(BLOCK), 0, s, keys, vals, flags, n, ht, len, min);
    check_launch("build_keyvals");
}

void launch_probe_semi(const int32_t *probe, const uint8_t *in, uint8_t *out, int64_t n, const uint8_t *ht,
                       int32_t min, int32_t max, hipStream_t s)
{
    if (n <= 0) return;
    const int vec = aligned(probe, 16) && aligned(in, 4) && aligned(out, 4);
    hipLaunchKernelGGL(k_probe_semi, dim3(vec ? grid_rows(n) : grid_one(n)), dim3(BLOCK), 0, s, probe, in, out, n,
                       ht, min, max, vec);
    check_launch("probe_semi");
}

// a probe over at least this many rows, and at least this many rows per
// table slot, gathers from a 16-bit copy of a table of up to 2^26 slots
constexpr int64_t PROBE_U16_MIN_ROWS = 4 << 20;
constexpr int64_t PROBE_U16_MIN_RATIO = 16;
constexpr int64_t PROBE_U16_MAX_SLOTS = int64_t(1) << 26;

void launch_probe_full(const int32_t *probe, int32_t *val_out, const uint8_t *in, uint8_t *out, int64_t n,
                       const int32_t *ht, int32_t min, int32_t max, hipStream_t s)
{
    if (n <= 0) return;
    if (!aligned(ht, 8)) fail(SDB_ERR_INVALID_ARG, "key-value hash table must be 8-byte aligned");
    const int vec = aligned(probe, 16) && aligned(in, 4) && aligned(out, 4);
    PairsOrU16 t{reinterpret_cast<const int2 *>(ht), nullptr, nullptr};
    // a table much smaller than the probe column gets a 16-bit copy first
    // (two passes over the table; the probe then gathers 2 bytes per row
    // instead of 8)
    const int64_t len = max >= min ? (int64_t)max - min + 1 : 0;
    if (n >= PROBE_U16_MIN_ROWS && len > 0 && len * PROBE_U16_MIN_RATIO <= n && len <= PROBE_U16_MAX_SLOTS) {
        uint8_t *scr = static_cast<uint8_t *>(scratch_buffer(256 + (size_t)len * 2, s));
        int32_t *info = reinterpret_cast<int32_t *>(scr);
        uint16_t *u16 = reinterpret_cast<uint16_t *>(scr + 256);
        const unsigned g = (unsigned)std::min<int64_t>((len + BLOCK - 1) / BLOCK, 1024);
        hipLaunchKernelGGL(k_pairs_range_init, dim3(1), dim3(1), 0, s, info);
        hipLaunchKernelGGL(k_pairs_range, dim3(g), dim3(BLOCK), 0, s, t.pairs, len, info);
        hipLaunchKernelGGL(k_pairs_to_u16, dim3(g), dim3(BLOCK), 0, s, t.pairs, len, info, u16);
        t.u16 = u16;
        t.info = info;
    }
    hipLaunchKernelGGL(k_probe_full, dim3(vec ? grid_rows(n) : grid_one(n)), dim3(BLOCK), 0, s, probe, val_out, in,
                       out, n, t, min, max, vec);
    check_launch("probe_full");
}

void launch_sum_flagged(const int32_t *col, const uint8_t *flags, int64_t n, uint64_t *acc, hipStream_t s)
{
    if (n <= 0) return;
    if (!aligned(acc, 8)) fail(SDB_ERR_INVALID_ARG, "accumulator must be 8-byte aligned");
    const int vec = aligned(col, 16) && aligned(flags, 4);
    hipLaunchKernelGGL(k_sum_flagged, dim3(vec ? grid_rows(n) : grid_one(n)), dim3(BLOCK), 0, s, col, flags, n, acc,
                       vec);
    check_launch("sum_flagged");
}

void launch_count_flags(const uint8_t *flags, int64_t n, uint64_t *out, bool accumulate, hipStream_t s)
{
    if (!aligned(out, 8)) fail(SDB_ERR_INVALID_ARG, "counter must be 8-byte aligned");
    if (!accumulate) {
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, out);
        check_launch("count_flags init");
    }
    if (n <= 0) return;
    const int vec = aligned(flags, 4);
    hipLaunchKernelGGL(k_count_flags, dim3(stream_grid(n, BLOCK * 4)), dim3(BLOCK), 0, s, flags, n, out, vec);
    check_launch("count_flags");
}

// LDS budget of the privatised group-by (bytes); the slot array is u64 plus a
// one-bit occupancy map.
constexpr uint64_t GROUPBY_LDS_BYTES = 64 * 1024;
// LDS of one CU (gfx950)
constexpr size_t GROUPBY_LDS_PER_CU = 160 * 1024;

uint64_t groupby_lds_max_slots() { return (GROUPBY_LDS_BYTES * 8) / (64 + 1); }

void launch_groupby_dense(const int32_t *const *keys, int k, const int32_t *mins, const int32_t *maxs,
                          const int32_t *agg, const uint8_t *flags, int64_t n, int32_t *const *out_keys,
                          uint64_t *sums, uint32_t *seen, uint64_t prod, hipStream_t s, int strategy)
{
    if (k < 1 || k > GB_MAX_KEYS) fail(SDB_ERR_INVALID_ARG, "group-by supports 1..8 key columns");
    if (prod == 0) fail(SDB_ERR_INVALID_ARG, "group-by range product is 0");
    if (n <= 0) return;
    GroupArgs g{};
    g.k = k;
    for (int j = 0; j < k; j++) {
        g.keys[j] = keys[j];
        g.out_keys[j] = out_keys[j];
        g.mins[j] = mins[j];
        g.maxs[j] = maxs[j];
        g.range[j] = (uint32_t)maxs[j] - (uint32_t)mins[j] + 1u;
    }
    // the slot decode at flush is exact when the in-range index cannot wrap
    uint64_t exact = 1;
    bool fits = true;
    for (int j = 0; j < k; j++) {
        if (maxs[j] < mins[j]) fits = false;
        exact *= (uint64_t)((int64_t)maxs[j] - (int64_t)mins[j] + 1);
        if (exact > prod) fits = false;
    }
    fits = fits && exact == prod;
    const bool lds = strategy != 1 && fits && prod <= groupby_lds_max_slots();
    if (lds) {
        // 1024-thread blocks share one LDS copy of the sums among 16 waves;
        // as many blocks per CU as that LDS allows (at most 2)
        const size_t bytes = prod * 8 + ((prod + 31) / 32) * 4;
        const int per_cu = std::max(1, std::min(2, (int)(GROUPBY_LDS_PER_CU / std::max<size_t>(bytes, 1))));
        const int vec = ((uintptr_t)flags % 4) == 0;
        const int64_t units = vec ? (n + 3) / 4 : n;
        hipLaunchKernelGGL(k_groupby_lds, dim3(stream_grid(units, (int64_t)GB_BLOCK * (vec ? UNROLL : 1), per_cu)),
                           dim3(GB_BLOCK), bytes, s, g, agg, flags, n, sums, seen, (uint32_t)prod, vec);
    } else {
        hipLaunchKernelGGL(k_groupby_global, dim3(grid_one(n)), dim3(BLOCK), 0, s, g, agg, flags, n, sums, seen,
                           prod);
    }
    check_launch("groupby_dense");
}

void launch_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod, hipStream_t s)
{
    if (prod == 0) return;
    hipLaunchKernelGGL(k_seen_to_flags, dim3(grid_one((int64_t)prod)), dim3(BLOCK), 0, s, seen, flags, prod);
    check_launch("seen_to_flags");
}

uint32_t bits_partials(int32_t len)
{
    if (len <= 0) return 0;
    const int64_t per_block = 64 * BITS_GROUPS * (BLOCK / 64);
    return (uint32_t)std::min<int64_t>(BITS_MAX_BLOCKS, ((int64_t)len + per_block - 1) / per_block);
}

void launch_bits_from_keys(const uint8_t *ht, int32_t len, uint32_t *bits, uint32_t *partials, hipStream_t s)
{
    if (len <= 0) return;
    hipLaunchKernelGGL(k_bits_from_keys, dim3(bits_partials(len)), dim3(BLOCK), 0, s, ht, len, bits, partials);
    check_launch("bits_from_keys");
}

void launch_bits_from_pairs(const int32_t *ht, int32_t len, uint32_t *bits, uint16_t *vals, int32_t vmin,
                            uint32_t *partials, uint32_t *bad, hipStream_t s)
{
    if (len <= 0) return;
    if (!aligned(ht, 8)) fail(SDB_ERR_INVALID_ARG, "key-value hash table must be 8-byte aligned");
    hipLaunchKernelGGL(k_bits_from_pairs, dim3(bits_partials(len)), dim3(BLOCK), 0, s,
                       reinterpret_cast<const int2 *>(ht), len, bits, vals, vmin, partials, bad);
    check_launch("bits_from_pairs");
}

void launch_stats_reduce(const StatsParts &p, uint32_t *dst, hipStream_t s)
{
    if (p.n <= 0) return;
    if (p.n > PACK_MAX_PARTS) fail(SDB_ERR_INVALID_ARG, "stats reduce: too many tables");
    hipLaunchKernelGGL(k_stats_reduce, dim3(1, (unsigned)p.n), dim3(BLOCK), 0, s, p, dst);
    check_launch("stats_reduce");
}

void launch_rank_dict(const uint32_t *bits, int32_t len, const uint16_t *vals, const uint32_t *partials,
                      uint8_t *dict, uint32_t dict_cap, hipStream_t s)
{
    if (len <= 0) return;
    const int32_t words = (int32_t)(((int64_t)len + 31) / 32);
    if (dict_cap < rank_values_off((uint32_t)words)) fail(SDB_ERR_RUNTIME, "rank dictionary buffer too small");
    const uint32_t cap_vals = (dict_cap - rank_values_off((uint32_t)words)) / 2u;
    // the bitmap kernel's blocks each counted one run of 32 words when its
    // grid covers the table without wrapping (below 2^21 slots; dictionaries
    // are built for tables far smaller)
    static_assert(64 * BITS_GROUPS * (BLOCK / 64) == 32 * RANK_PART_WORDS, "one bitmap block per 32 words");
    const uint32_t ranges = (uint32_t)(((int64_t)words + RANK_PART_WORDS - 1) / RANK_PART_WORDS);
    if (!partials || ranges != bits_partials(len))
        fail(SDB_ERR_RUNTIME, "rank dictionary: the bitmap's per-block counts do not cover 32 words each");
    hipLaunchKernelGGL(k_rank_dict_parts, dim3(ranges), dim3(64), 0, s, bits, words, partials, vals, (uint32_t)len,
                       dict, cap_vals);
    check_launch("rank_dict");
}

void launch_batch_masks(const FMaskCol *cols, int ncols, uint32_t max_len, const FMaskEntry *ents, hipStream_t s)
{
    if (ncols <= 0 || max_len == 0) return;
    const unsigned gx = (unsigned)std::min<uint64_t>(((uint64_t)max_len + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_batch_masks, dim3(gx, (unsigned)ncols), dim3(BLOCK), 0, s, cols, ents);
    check_launch("batch_masks");
}

void launch_batch_dict(const FMaskDict *dicts, int ndicts, hipStream_t s)
{
    if (ndicts <= 0) return;
    hipLaunchKernelGGL(k_batch_dict, dim3((unsigned)ndicts), dim3(DICT_BLOCK), 0, s, dicts);
    check_launch("batch_dict");
}

int64_t compact_tiles(int64_t n) { return n > 0 ? (n + COMPACT_TILE - 1) / COMPACT_TILE : 0; }

void launch_compact(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, hipStream_t s)
{
    const int64_t ntiles = compact_tiles(n);
    uint32_t *counts = reinterpret_cast<uint32_t *>(scratch_buffer((size_t)std::max<int64_t>(ntiles, 1) * 4, s));
    launch_compact_into(flags, n, row_ids, count, counts, s);
}

void launch_compact_into(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, uint32_t *counts,
                         hipStream_t s)
{
    const int64_t ntiles = compact_tiles(n);
    if (ntiles > 0)
        hipLaunchKernelGGL(k_compact_count, dim3((unsigned)ntiles), dim3(BLOCK), 0, s, flags, n, counts);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(BLOCK), 0, s, counts, ntiles, count);
    if (ntiles > 0)
        hipLaunchKernelGGL(k_compact_write, dim3((unsigned)ntiles), dim3(BLOCK), 0, s, flags, n, counts, row_ids);
    check_launch("compact");
}

// ------------------------------------------- compaction without allocation
//
// The per-tile counts of sdb_compact_noalloc live in a fixed device array of
// the library instead of library-allocated scratch: each GPU the library runs
// on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel (on the host it only stands for the symbol); its
// device address is looked up once per device, with that device current, by
// hipGetSymbolAddress, and that address is what the kernels get.  Every call
// on one GPU shares the array; the counting pass overwrites every count the
// scan and write passes read, so nothing needs clearing between calls.
constexpr int64_t NOALLOC_TILES = 65536; // 256 KiB of counts: 2^28 rows
static_assert(NOALLOC_TILES * COMPACT_TILE == (int64_t(1) << 28), "noalloc compaction covers 2^28 rows");
static __device__ __align__(256) uint32_t scratch[NOALLOC_TILES];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
uint32_t *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

uint32_t *noalloc_counts()
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) fail(SDB_ERR_UNSUPPORTED, "device ordinal beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) fail(SDB_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_noalloc_ptr[dev] = static_cast<uint32_t *>(ptr);
    }
    return g_noalloc_ptr[dev];
}

void launch_compact_noalloc(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, hipStream_t s)
{
    if (n > NOALLOC_TILES * COMPACT_TILE)
        fail(SDB_ERR_INVALID_ARG, "sdb_compact_noalloc: n exceeds 2^28 rows (the fixed tile-count array)");
    launch_compact_into(flags, n, row_ids, count, noalloc_counts(), s);
}

void launch_gather_cols(const GatherCols &g, const int32_t *ids, int64_t count, hipStream_t s)
{
    if (count <= 0 || g.ncols <= 0) return;
    if (g.ncols > GATHER_MAX_COLS) fail(SDB_ERR_INVALID_ARG, "gather: too many columns");
    hipLaunchKernelGGL(k_gather_cols, dim3(grid_one(count)), dim3(BLOCK), 0, s, g, ids, count);
    check_launch("gather_cols");
}

void launch_pack_parts(const PackParts &p, uint8_t *dst, hipStream_t s)
{
    if (p.n <= 0) return;
    if (p.n > PACK_MAX_PARTS) fail(SDB_ERR_INVALID_ARG, "pack: too many parts");
    uint64_t most = 0;
    for (int i = 0; i < p.n; i++) most = std::max(most, p.bytes[i]);
    const unsigned gx = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((most / 16 + BLOCK - 1) / BLOCK, 256));
    hipLaunchKernelGGL(k_pack_parts, dim3(gx, (unsigned)p.n), dim3(BLOCK), 0, s, p, dst);
    check_launch("pack_parts");
}

void launch_gather_i32(const int32_t *src, const int32_t *ids, int64_t count, int32_t *dst, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_gather_i32, dim3(grid_one(count)), dim3(BLOCK), 0, s, src, ids, count, dst);
    check_launch("gather_i32");
}

void minmax_i32(const int32_t *col, int64_t n, int32_t *mn, int32_t *mx, hipStream_t s)
{
    if (n <= 0) fail(SDB_ERR_INVALID_ARG, "min/max of an empty column");
    int32_t *mm = reinterpret_cast<int32_t *>(scratch_buffer(8, s));
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, s, mm);
    hipLaunchKernelGGL(k_minmax, dim3(grid_one(n)), dim3(BLOCK), 0, s, col, n, mm);
    check_launch("minmax");
    int32_t host[2];
    SDB_HIP(hipMemcpyAsync(host, mm, 8, hipMemcpyDeviceToHost, s));
    SDB_HIP(hipStreamSynchronize(s));
    *mn = host[0];
    *mx = host[1];
}

} // namespace sdb

// =================================================================== C ABI

using namespace sdb;

namespace {
inline void check_n(int64_t n) { require(n >= 0, "negative length"); }
inline void check_cmp(int cmp) { require(cmp >= SDB_EQ && cmp <= SDB_GE, "comparison op out of range"); }
inline void check_logic(int l) { require(l >= SDB_NONE && l <= SDB_OR, "logical op out of range"); }
inline void check_binop(int op) { require(op >= SDB_MUL && op <= SDB_SUB, "arithmetic op out of range"); }
} // namespace

extern "C" {

int sdb_comp_op_from_string(const char *op)
{
    const std::string s = op ? op : "";
    if (s == "==") return SDB_EQ;
    if (s == ">=") return SDB_GE;
    if (s == "<=") return SDB_LE;
    if (s == "<") return SDB_LT;
    if (s == ">") return SDB_GT;
    if (s == "!=" || s == "<>") return SDB_NE;
    return SDB_EQ;
}

int sdb_logical_op_from_string(const char *op)
{
    const std::string s = op ? op : "";
    if (s == "AND") return SDB_AND;
    if (s == "OR") return SDB_OR;
    return SDB_NONE;
}

int sdb_binop_from_string(const char *op)
{
    const std::string s = op ? op : "";
    if (s == "*") return SDB_MUL;
    if (s == "/") return SDB_DIV;
    if (s == "+") return SDB_ADD;
    if (s == "-") return SDB_SUB;
    return -1;
}

int sdb_select_lit(const int32_t *col, const uint8_t *in_flags, uint8_t *out_flags, int64_t n, int cmp, int logic,
                   int32_t value, sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_cmp(cmp), check_logic(logic);
        require(n == 0 || (col && in_flags && out_flags), "null pointer");
        launch_select_lit(col, in_flags, out_flags, n, cmp, logic, value, as_stream(stream));
    });
}

int sdb_select_col(const int32_t *a, const int32_t *b, const uint8_t *in_flags, uint8_t *out_flags, int64_t n,
                   int cmp, int logic, sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_cmp(cmp), check_logic(logic);
        require(n == 0 || (a && b && in_flags && out_flags), "null pointer");
        launch_select_col(a, b, in_flags, out_flags, n, cmp, logic, as_stream(stream));
    });
}

int sdb_logical(const uint8_t *in_flags, const uint8_t *local_flags, uint8_t *out_flags, int64_t n, int logic,
                sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_logic(logic);
        require(n == 0 || (in_flags && local_flags && out_flags), "null pointer");
        launch_logical(in_flags, local_flags, out_flags, n, logic, as_stream(stream));
    });
}

int sdb_fill_i32(int32_t *dst, int32_t value, int64_t n, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(n == 0 || dst, "null pointer");
        require(((uintptr_t)dst & 3) == 0, "int32 column must be 4-byte aligned");
        launch_fill_i32(dst, value, n, as_stream(stream));
    });
}

int sdb_binop_cc(int32_t *result, const int32_t *a, const int32_t *b, const uint8_t *flags, int64_t n, int op,
                 sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_binop(op);
        require(n == 0 || (result && a && b && flags), "null pointer");
        launch_binop(0, op, result, a, b, 0, flags, n, as_stream(stream));
    });
}

int sdb_binop_lc(int32_t *result, int32_t a, const int32_t *b, const uint8_t *flags, int64_t n, int op,
                 sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_binop(op);
        require(n == 0 || (result && b && flags), "null pointer");
        launch_binop(1, op, result, nullptr, b, a, flags, n, as_stream(stream));
    });
}

int sdb_binop_cl(int32_t *result, const int32_t *a, int32_t b, const uint8_t *flags, int64_t n, int op,
                 sdb_stream stream)
{
    return guarded([&] {
        check_n(n), check_binop(op);
        require(n == 0 || (result && a && flags), "null pointer");
        launch_binop(2, op, result, a, nullptr, b, flags, n, as_stream(stream));
    });
}

int sdb_build_keys(const int32_t *keys, const uint8_t *flags, int64_t n, uint8_t *ht, int32_t ht_len,
                   int32_t ht_min, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(ht_len > 0, "hash table length must be positive");
        require(n == 0 || (keys && flags && ht), "null pointer");
        launch_build_keys(keys, flags, n, ht, ht_len, ht_min, as_stream(stream));
    });
}

int sdb_build_keyvals(const int32_t *keys, const int32_t *vals, const uint8_t *flags, int64_t n, int32_t *ht,
                      int32_t ht_len, int32_t ht_min, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(ht_len > 0, "hash table length must be positive");
        require(n == 0 || (keys && vals && flags && ht), "null pointer");
        launch_build_keyvals(keys, vals, flags, n, ht, ht_len, ht_min, as_stream(stream));
    });
}

int sdb_probe_semi(const int32_t *probe, const uint8_t *in_flags, uint8_t *out_flags, int64_t n, const uint8_t *ht,
                   int32_t ht_min, int32_t ht_max, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(ht_max >= ht_min, "ht_max < ht_min");
        require(n == 0 || (probe && in_flags && out_flags && ht), "null pointer");
        launch_probe_semi(probe, in_flags, out_flags, n, ht, ht_min, ht_max, as_stream(stream));
    });
}

int sdb_probe_full(const int32_t *probe, int32_t *val_out, const uint8_t *in_flags, uint8_t *out_flags, int64_t n,
                   const int32_t *ht, int32_t ht_min, int32_t ht_max, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(ht_max >= ht_min, "ht_max < ht_min");
        require(n == 0 || (probe && val_out && in_flags && out_flags && ht), "null pointer");
        launch_probe_full(probe, val_out, in_flags, out_flags, n, ht, ht_min, ht_max, as_stream(stream));
    });
}

int sdb_sum_flagged(const int32_t *col, const uint8_t *flags, int64_t n, uint64_t *acc, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(acc, "null accumulator");
        require(n == 0 || (col && flags), "null pointer");
        launch_sum_flagged(col, flags, n, acc, as_stream(stream));
    });
}

int sdb_groupby_dense(const int32_t *const *keys, int k, const int32_t *mins, const int32_t *maxs,
                      const int32_t *agg, const uint8_t *flags, int64_t n, int32_t *const *out_keys,
                      uint64_t *sums, uint32_t *seen, uint64_t prod, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(keys && mins && maxs && out_keys, "null pointer");
        require(prod == 0 || (sums && seen), "null pointer");
        require(n == 0 || (agg && flags), "null pointer");
        for (int j = 0; j < k && j < 8; j++) require(n == 0 || (keys[j] && out_keys[j]), "null key column");
        launch_groupby_dense(keys, k, mins, maxs, agg, flags, n, out_keys, sums, seen, prod, as_stream(stream), 0);
    });
}

int sdb_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod, sdb_stream stream)
{
    return guarded([&] {
        require(prod == 0 || (seen && flags), "null pointer");
        launch_seen_to_flags(seen, flags, prod, as_stream(stream));
    });
}

int sdb_count_flags(const uint8_t *flags, int64_t n, uint64_t *out, int accumulate, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(out, "null output");
        require(n == 0 || flags, "null pointer");
        launch_count_flags(flags, n, out, accumulate != 0, as_stream(stream));
    });
}

int sdb_compact(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(n <= INT32_MAX, "compaction is limited to 2^31-1 rows (int32 row ids)");
        require(count, "null count");
        require(n == 0 || (flags && row_ids), "null pointer");
        launch_compact(flags, n, row_ids, count, as_stream(stream));
    });
}

int sdb_compact_noalloc(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(n <= SDB_COMPACT_NOALLOC_MAX_ROWS, "sdb_compact_noalloc: n exceeds SDB_COMPACT_NOALLOC_MAX_ROWS (2^28)");
        require(count, "null count");
        require(n == 0 || (flags && row_ids), "null pointer");
        launch_compact_noalloc(flags, n, row_ids, count, as_stream(stream));
    });
}

int sdb_gather_i32(const int32_t *src, const int32_t *row_ids, int64_t count, int32_t *dst, sdb_stream stream)
{
    return guarded([&] {
        check_n(count);
        require(count == 0 || (src && row_ids && dst), "null pointer");
        launch_gather_i32(src, row_ids, count, dst, as_stream(stream));
    });
}

int sdb_compress_sync(const int32_t *const *cols, int ncols, const uint8_t *flags, int64_t n,
                      int32_t *const *host_cols, uint64_t *count, sdb_stream stream)
{
    return guarded([&] {
        check_n(n);
        require(n <= INT32_MAX, "compress_sync is limited to 2^31-1 rows (int32 row ids)");
        require(ncols >= 0 && ncols <= GATHER_MAX_COLS, "compress_sync: 0 to 8 columns");
        require(count && (ncols == 0 || (cols && host_cols)), "null pointer");
        *count = 0;
        if (n == 0) return;
        for (int c = 0; c < ncols; c++) require(cols[c] && host_cols[c], "null column");
        require(flags, "null flags");
        hipStream_t s = as_stream(stream);
        // device half: the selected rows' ids (build_row_ids,
        // models/transient_table.hpp:535-599, here in row order) ...
        const int64_t ntiles = compact_tiles(n);
        const size_t ids_bytes = ((size_t)n * 4 + 255) / 256 * 256, cnt_off = ids_bytes + (size_t)ntiles * 4;
        uint8_t *scr = static_cast<uint8_t *>(scratch_buffer(cnt_off + 256, s));
        int32_t *ids = reinterpret_cast<int32_t *>(scr);
        uint32_t *dcount = reinterpret_cast<uint32_t *>(scr + cnt_off);
        launch_compact_into(flags, n, ids, dcount, reinterpret_cast<uint32_t *>(scr + ids_bytes), s);
        uint32_t selected = 0;
        SDB_HIP(hipMemcpyAsync(&selected, dcount, 4, hipMemcpyDeviceToHost, s));
        SDB_HIP(hipStreamSynchronize(s));
        *count = selected;
        if (selected == 0 || ncols == 0) return;
        // ... and their values, gathered (Segment::compress_sync,
        // models/models.hpp:1019-1035) and copied to the host with the ids
        int32_t *vals = nullptr;
        SDB_HIP(hipMalloc(&vals, (size_t)selected * 4 * (size_t)ncols));
        std::vector<int32_t> hids(selected), hvals((size_t)selected * (size_t)ncols);
        hipError_t e = hipSuccess;
        try {
            GatherCols g{};
            for (int c = 0; c < ncols; c++) {
                g.src[c] = cols[c];
                g.dst[c] = vals + (size_t)c * selected;
                g.src_rows[c] = (uint64_t)n;
            }
            g.ncols = ncols;
            launch_gather_cols(g, ids, selected, s);
            e = hipMemcpyAsync(hids.data(), ids, (size_t)selected * 4, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hvals.data(), vals, hvals.size() * 4, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(vals);
            throw;
        }
        (void)hipFree(vals);
        hip_check(e, "compress_sync copy");
        // host half: the scatter into the host columns (models.hpp:1044-1063)
        for (int c = 0; c < ncols; c++) {
            const int32_t *v = hvals.data() + (size_t)c * selected;
            int32_t *dst = host_cols[c];
            for (uint32_t j = 0; j < selected; j++) dst[hids[j]] = v[j];
        }
    });
}

int sdb_minmax_i32(const int32_t *col, int64_t n, int32_t *min_out, int32_t *max_out, sdb_stream stream)
{
    return guarded([&] {
        require(n > 0, "min/max needs at least one row");
        require(col && min_out && max_out, "null pointer");
        minmax_i32(col, n, min_out, max_out, as_stream(stream));
    });
}

} // extern "C"
