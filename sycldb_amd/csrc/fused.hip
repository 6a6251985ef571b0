// The fused bundle kernel: one launch runs a whole per-segment operator
// sequence over the segment's rows (see fused.hpp for the program format).
//
// Each thread executes the program on chunks of R rows (R = 4: 16-byte column
// loads; R = 1: any alignment / the tail).  Flag registers of a chunk live in
// one 32-bit word (4 bits per register); value registers live in an LDS
// register file (one int4 per register and thread), addressed by the uniform
// register number.  The instruction stream is uniform across the grid, so the
// dispatch on the instruction kind is a scalar branch; per-row work follows the
// reference functors exactly.  A column operand is only loaded for chunks in
// which some row still needs it, so a selective prefix of the program cuts the
// HBM lines later columns pull.  Small probe bitmaps are staged in LDS at kernel
// start.  Terminal reductions are privatised per thread (sum) or per block in
// LDS (dense slots or an open-addressed hash of slots) and flushed once.
#include "capi_util.hpp"
#include "fused.hpp"
#include "fused_device.hpp"

#include <algorithm>
#include <map>
#include <mutex>

namespace sdb {
namespace {

using namespace fdev;

// Value registers: register j of thread t is regs[j * BLOCK + t] (an int4 =
// the rows of one chunk).  The register number is uniform, so an access is
// one ds_read/ds_write_b128 at a computed address.
struct VRegs {
    int4 *base; // &regs[threadIdx.x]
};

template <int R> __device__ __forceinline__ void vget(const VRegs &V, int r, int32_t (&out)[R])
{
    const int4 t = V.base[r * BLOCK];
    out[0] = t.x;
    if constexpr (R == 4) {
        out[1] = t.y;
        out[2] = t.z;
        out[3] = t.w;
    }
}

template <int R> __device__ __forceinline__ void vset(const VRegs &V, int r, const int32_t (&in)[R], uint32_t mask)
{
    constexpr uint32_t M = (1u << R) - 1;
    if constexpr (R == 4) {
        if ((mask & M) == M) {
            V.base[r * BLOCK] = make_int4(in[0], in[1], in[2], in[3]);
            return;
        }
        int4 t = V.base[r * BLOCK];
        if (mask & 1u) t.x = in[0];
        if (mask & 2u) t.y = in[1];
        if (mask & 4u) t.z = in[2];
        if (mask & 8u) t.w = in[3];
        V.base[r * BLOCK] = t;
    } else {
        if (mask & 1u) V.base[r * BLOCK].x = in[0];
    }
}

// Prefetched column values of the chunk being run: row k of c<j> is
// pf_mem[j][row0 + k].  Four named vector values, not an array: a register
// picked by a uniform number out of an array would become a dynamically
// indexed scratch access.
template <int R> struct PfType { typedef v4i type; };
template <> struct PfType<1> { typedef int32_t type; };
template <int R> struct Pf {
    typename PfType<R>::type c0, c1, c2, c3;
    __device__ __forceinline__ typename PfType<R>::type pick(int j) const
    {
        return j == 0 ? c0 : j == 1 ? c1 : j == 2 ? c2 : c3;
    }
};
static_assert(F_MAX_PF == 4, "Pf holds four columns");

template <int R> __device__ __forceinline__ typename PfType<R>::type pf_col(const int32_t *mem, int64_t row0)
{
    if constexpr (R == 4) return __builtin_nontemporal_load(reinterpret_cast<const v4i *>(mem + row0));
    else return __builtin_nontemporal_load(mem + row0);
}

template <int R> __device__ __forceinline__ void pf_load(const FProgram &P, int64_t row0, Pf<R> &pf)
{
    if (P.npf > 0) pf.c0 = pf_col<R>(P.pf_mem[0], row0);
    if (P.npf > 1) pf.c1 = pf_col<R>(P.pf_mem[1], row0);
    if (P.npf > 2) pf.c2 = pf_col<R>(P.pf_mem[2], row0);
    if (P.npf > 3) pf.c3 = pf_col<R>(P.pf_mem[3], row0);
}

// Column operand: a prefetched column, a value register, or R rows of a
// column loaded now.  Column reads are streaming (non-temporal) so they do not
// push probe tables out of the caches.  Register numbers are uniform, so the
// prefetch slot is picked by scalar compares.
template <int R>
__device__ __forceinline__ void operand(int reg, const int32_t *mem, int64_t row0, const VRegs &V, const Pf<R> &pf,
                                        int32_t (&x)[R])
{
    if (reg >= F_PF_BASE) {
        const auto t = pf.pick(reg - F_PF_BASE);
        if constexpr (R == 4) {
            x[0] = t.x;
            x[1] = t.y;
            x[2] = t.z;
            x[3] = t.w;
        } else {
            x[0] = t;
        }
    } else if (reg >= 0) {
        vget<R>(V, reg, x);
    } else if constexpr (R == 4) {
        const v4i t = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(mem + row0));
        x[0] = t.x;
        x[1] = t.y;
        x[2] = t.z;
        x[3] = t.w;
    } else {
        x[0] = __builtin_nontemporal_load(mem + row0);
    }
}

// Row k of an operand (see operand()).
template <int R>
__device__ __forceinline__ int32_t operand_row(int reg, const int32_t *mem, int64_t row0, int k, const VRegs &V,
                                               const Pf<R> &pf)
{
    if (reg >= F_PF_BASE) {
        const auto t = pf.pick(reg - F_PF_BASE);
        if constexpr (R == 4) return t[k];
        else return t;
    }
    if (reg >= 0) return reinterpret_cast<const int32_t *>(&V.base[reg * BLOCK])[k];
    return __builtin_nontemporal_load(mem + row0 + k);
}

template <int R> __device__ __forceinline__ void store_masked(int32_t *dst, int64_t row0, const int32_t (&x)[R], uint32_t m)
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

struct Lds {
    uint8_t *base; // dynamic LDS
};

template <int R>
__device__ __forceinline__ void run_chunk(const FProgram &P, int64_t row0, int64_t &acc, const Lds &L,
                                          const VRegs &V, const Pf<R> &pf)
{
    constexpr uint32_t M = (1u << R) - 1;
    uint32_t F = 0;
#pragma unroll
    for (int j = 0; j < F_MAX_FLAGS; j++) {
        const uint8_t init = P.flag_init[j];
        if (init == FF_LOAD && P.flag_load[j]) F |= flags_in<R>(P.flag_load[j] + row0) << (4 * j);
        else if (init == FF_TRUE) F |= M << (4 * j);
    }

    for (int i = 0; i < P.ninstr; i++) {
        const FInstr &I = P.ins[i];
        const uint32_t in = (F >> (4 * I.fin)) & M;
        switch (I.kind) {
        case FK_SEL_LIT:
        case FK_SEL_COL:
        case FK_SEL_RANGE:
        case FK_SEL_IN2: {
            const uint32_t need = I.logic == SDB_NONE ? M : (I.logic == SDB_AND ? in : (~in & M));
            uint32_t cm = 0;
            if (need) {
                int32_t x[R], y[R];
                operand<R>(I.va, I.a, row0, V, pf, x);
                if (I.kind == FK_SEL_COL) operand<R>(I.vb, I.b, row0, V, pf, y);
#pragma unroll
                for (int k = 0; k < R; k++) {
                    bool c;
                    if (I.kind == FK_SEL_RANGE) c = x[k] >= I.lit && x[k] <= I.lit2;
                    else if (I.kind == FK_SEL_IN2) c = x[k] == I.lit || x[k] == I.lit2;
                    else c = compare(I.cmp, x[k], I.kind == FK_SEL_COL ? y[k] : I.lit);
                    cm |= (uint32_t)c << k;
                }
            }
            F = (F & ~(0xFu << (4 * I.fout))) | (logic_combine(I.logic, in, cm) << (4 * I.fout));
            break;
        }
        case FK_LOGICAL: {
            const uint32_t loc = (F >> (4 * I.floc)) & M;
            F = (F & ~(0xFu << (4 * I.fout))) | (logic_combine(I.logic, in, loc) << (4 * I.fout));
            break;
        }
        case FK_FILL: {
            int32_t x[R];
#pragma unroll
            for (int k = 0; k < R; k++) x[k] = I.lit;
            vset<R>(V, I.vout, x, M);
            if (I.store) store_masked<R>(I.store, row0, x, M);
            break;
        }
        case FK_BINOP_CC:
        case FK_BINOP_LC:
        case FK_BINOP_CL: {
            if (!in) break;
            int32_t x[R], y[R], r[R];
            if (I.kind != FK_BINOP_LC) operand<R>(I.va, I.a, row0, V, pf, x);
            if (I.kind != FK_BINOP_CL) operand<R>(I.vb, I.b, row0, V, pf, y);
#pragma unroll
            for (int k = 0; k < R; k++)
                r[k] = element_op(I.op, I.kind == FK_BINOP_LC ? I.lit : x[k], I.kind == FK_BINOP_CL ? I.lit : y[k]);
            vset<R>(V, I.vout, r, in);
            if (I.store) store_masked<R>(I.store, row0, r, in);
            break;
        }
        case FK_PROBE_SEMI: {
            uint32_t out = 0;
            if (in) {
                int32_t x[R];
                operand<R>(I.va, I.a, row0, V, pf, x);
                const uint8_t *ht = reinterpret_cast<const uint8_t *>(I.ht);
#pragma unroll
                for (int k = 0; k < R; k++)
                    if (((in >> k) & 1u) && x[k] >= I.ht_min && x[k] <= I.ht_max && ht[slot_of(x[k], I.ht_min)] != 0)
                        out |= 1u << k;
            }
            F = (F & ~(0xFu << (4 * I.fout))) | (out << (4 * I.fout));
            break;
        }
        case FK_PROBE_FULL: {
            uint32_t out = 0;
            if (in) {
                int32_t x[R], r[R];
                operand<R>(I.va, I.a, row0, V, pf, x);
                const int2 *ht = reinterpret_cast<const int2 *>(I.ht);
#pragma unroll
                for (int k = 0; k < R; k++) {
                    r[k] = 0;
                    if (((in >> k) & 1u) && x[k] >= I.ht_min && x[k] <= I.ht_max) {
                        const int2 e = ht[slot_of(x[k], I.ht_min)];
                        if (e.x == 1) {
                            r[k] = e.y;
                            out |= 1u << k;
                        }
                    }
                }
                vset<R>(V, I.vout, r, out);
                if (I.store && out) store_masked<R>(I.store, row0, r, out);
            }
            F = (F & ~(0xFu << (4 * I.fout))) | (out << (4 * I.fout));
            break;
        }
        case FK_PROBE_BITS: {
            // presence bit of a slot = the reference table's "match"; the value,
            // when wanted, comes from the compact table or the reference pairs
            uint32_t out = 0;
            if (in) {
                int32_t x[R];
                uint32_t slot[R];
                operand<R>(I.va, I.a, row0, V, pf, x);
                const uint32_t *bits = I.in_lds ? reinterpret_cast<const uint32_t *>(L.base + I.lds_off)
                                                : reinterpret_cast<const uint32_t *>(I.ht);
#pragma unroll
                for (int k = 0; k < R; k++) {
                    slot[k] = slot_of(x[k], I.ht_min);
                    if (((in >> k) & 1u) && x[k] >= I.ht_min && x[k] <= I.ht_max &&
                        ((bits[slot[k] >> 5] >> (slot[k] & 31)) & 1u))
                        out |= 1u << k;
                }
                if (I.vkind != FV_NONE && out) {
                    int32_t r[R];
#pragma unroll
                    for (int k = 0; k < R; k++) {
                        r[k] = 0;
                        if ((out >> k) & 1u) {
                            if (I.vkind == FV_U16) {
                                r[k] = (int32_t)((uint32_t)I.vmin + reinterpret_cast<const uint16_t *>(I.vals)[slot[k]]);
                            } else if (I.vkind == FV_RANK16) {
                                // rank dictionary in LDS: prefix count of the word + set bits below
                                const uint8_t *dict = L.base + I.lds_off;
                                const uint32_t w = slot[k] >> 5;
                                const uint32_t rank =
                                    reinterpret_cast<const uint16_t *>(dict + rank_prefix_off((uint32_t)I.ht_len))[w] +
                                    __popc(bits[w] & ((1u << (slot[k] & 31)) - 1u));
                                r[k] = (int32_t)((uint32_t)I.vmin + reinterpret_cast<const uint16_t *>(
                                                                        dict + rank_values_off((uint32_t)I.ht_len))[rank]);
                            } else {
                                r[k] = reinterpret_cast<const int2 *>(I.vals)[slot[k]].y;
                            }
                        }
                    }
                    vset<R>(V, I.vout, r, out);
                    if (I.store) store_masked<R>(I.store, row0, r, out);
                }
            }
            F = (F & ~(0xFu << (4 * I.fout))) | (out << (4 * I.fout));
            break;
        }
        case FK_BUILD_KEYS: {
            int32_t x[R];
            operand<R>(I.va, I.a, row0, V, pf, x);
            uint8_t *ht = reinterpret_cast<uint8_t *>(const_cast<void *>(I.ht));
#pragma unroll
            for (int k = 0; k < R; k++) {
                const uint32_t s = slot_of(x[k], I.ht_min);
                if (s < (uint32_t)I.ht_len) ht[s] = (in >> k) & 1u;
            }
            break;
        }
        case FK_BUILD_KEYVALS: {
            int32_t x[R], y[R];
            operand<R>(I.va, I.a, row0, V, pf, x);
            if (in) operand<R>(I.vb, I.b, row0, V, pf, y);
            int32_t *ht = reinterpret_cast<int32_t *>(const_cast<void *>(I.ht));
#pragma unroll
            for (int k = 0; k < R; k++) {
                const uint32_t s = slot_of(x[k], I.ht_min);
                if (s >= (uint32_t)I.ht_len) continue;
                if ((in >> k) & 1u)
                    reinterpret_cast<int2 *>(ht)[s] = make_int2(1, y[k]);
                else
                    ht[2 * (int64_t)s] = 0;
            }
            break;
        }
        case FK_SUM: {
            if (!in) break;
            int32_t x[R];
            operand<R>(I.va, I.a, row0, V, pf, x);
#pragma unroll
            for (int k = 0; k < R; k++)
                if ((in >> k) & 1u) acc += x[k];
            break;
        }
        case FK_GROUPBY: {
            if (!in) break;
            const FGroup &g = P.gb;
            // row by row: only one row's keys are live at a time
#pragma unroll
            for (int k = 0; k < R; k++) {
                if (!((in >> k) & 1u)) continue;
                int32_t kv[F_MAX_KEYS];
#pragma unroll
                for (int j = 0; j < F_MAX_KEYS; j++)
                    kv[j] = j < g.k ? operand_row<R>(g.key_reg[j], g.key_mem[j], row0, k, V, pf) : 0;
                gb_row(g, L.base, kv, operand_row<R>(g.agg_reg, g.agg_mem, row0, k, V, pf));
            }
            break;
        }
        default: break;
        }
    }
#pragma unroll
    for (int j = 0; j < F_MAX_FLAGS; j++)
        if (P.flag_store[j]) flags_out<R>(P.flag_store[j] + row0, (F >> (4 * j)) & M);
}

// Runs chunks c, c + stride, ... below nchunks (chunk c = rows [c*R, c*R+R)).
// The prefetched columns of the next chunk are loaded before the current one
// runs, so their latency overlaps the current chunk's work.
template <int R>
__device__ __forceinline__ void stream_chunks(const FProgram &P, int64_t c, int64_t nchunks, int64_t stride,
                                              int64_t &acc, const Lds &L, const VRegs &V)
{
    Pf<R> next;
    if (c < nchunks) pf_load<R>(P, c * R, next);
    while (c < nchunks) {
        const Pf<R> cur = next;
        const int64_t cn = c + stride;
        if (cn < nchunks) pf_load<R>(P, cn * R, next);
        run_chunk<R>(P, c * R, acc, L, V, cur);
        c = cn;
    }
}

// The program is the kernel's first argument; it is read in place from the
// kernel-argument segment (constant address space, scalar loads) instead of
// through the by-value parameter, whose dynamic indexing would force a copy
// to scratch.  The magic word guards the layout assumption.
__global__ void __launch_bounds__(BLOCK) k_fused(const FProgram P_unused, int64_t n)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const FProgram &P = *(const FProgram __attribute__((address_space(4))) *)__builtin_amdgcn_kernarg_segment_ptr();
#else
    const FProgram &P = P_unused;
#endif
    if (P.magic != F_MAGIC) return;
    extern __shared__ uint4 dyn_lds[];
    const Lds L{reinterpret_cast<uint8_t *>(dyn_lds)};
    const VRegs V{reinterpret_cast<int4 *>(L.base + P.vreg_off) + threadIdx.x};

    block_prologue(P, L.base);

    int64_t acc = 0;
    if (P.vec) {
        const int64_t nchunks = n >> 2;
        stream_chunks<4>(P, (int64_t)blockIdx.x * BLOCK + threadIdx.x, nchunks, (int64_t)gridDim.x * BLOCK, acc, L, V);
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x < (n & 3)) {
            Pf<1> pf;
            const int64_t row0 = (nchunks << 2) + threadIdx.x;
            pf_load<1>(P, row0, pf);
            run_chunk<1>(P, row0, acc, L, V, pf);
        }
    } else {
        stream_chunks<1>(P, (int64_t)blockIdx.x * BLOCK + threadIdx.x, n, (int64_t)gridDim.x * BLOCK, acc, L, V);
    }
    block_epilogue(P, L.base, acc);
}

} // namespace

// Resident blocks of k_fused per CU with `lds` bytes of dynamic LDS, from the
// runtime's occupancy calculator (per device, cached per LDS size).
static int fused_blocks_per_cu(uint32_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, int> cache;
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, lds});
    if (it != cache.end()) return it->second;
    int blocks = 0;
    SDB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_fused, BLOCK, lds));
    blocks = std::max(1, blocks);
    cache[{dev, lds}] = blocks;
    return blocks;
}

void launch_fused(const FProgram &p, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    if (p.magic != F_MAGIC) fail(SDB_ERR_RUNTIME, "fused program without magic");
    if (p.lds_bytes > F_LDS_BUDGET) fail(SDB_ERR_RUNTIME, "fused program exceeds its LDS budget");
    const int64_t units = p.vec ? (n + 3) / 4 : n;
    // exactly the blocks that are resident at once (registers and LDS): the
    // grid-stride loop gives every block the same share, so a grid larger
    // than one resident wave of blocks would run its tail as a second wave
    hipLaunchKernelGGL(k_fused, dim3(stream_grid(units, BLOCK, fused_blocks_per_cu(p.lds_bytes))), dim3(BLOCK),
                       p.lds_bytes, s, p, n);
    hip_check(hipGetLastError(), "fused bundle");
}

} // namespace sdb
