// Program format of the fused execution of one KernelBundle (the per-segment
// operator sequence the reference queues in models/execution.hpp:500-651) as
// a single kernel.
//
// The reference designed this fusion (the JIT call sequence of
// kernels/jit_kernels.hpp:12-39 and models/execution.hpp:546-615) but runs
// with it disabled.  Here the host translates the bundle's operator list into
// a small program: every distinct flag buffer the bundle touches becomes a
// flag register, every column the bundle itself materialises becomes a value
// register, and each operator becomes one instruction with the reference
// functor's exact per-row semantics.  Per row the program therefore computes
// what the unfused sequence computes; only the stores differ: a buffer is
// written back only when it is still referenced after the bundle.
//
// Shared by the host, the compiled interpreter kernel and the kernels compiled
// at run time for one program shape (which see this file as an in-memory
// header), so all three agree on the layout.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sdb {

constexpr int F_MAX_INSTR = 24;
constexpr int F_MAX_FLAGS = 8;
constexpr int F_MAX_VALS = 4;
constexpr int F_MAX_KEYS = 4;
// Prefetched columns: operand register F_PF_BASE + j names column pf_mem[j],
// which the kernel loads for the next chunk while it runs the current one.
constexpr int F_MAX_PF = 4;
constexpr int F_PF_BASE = 16;

enum FKind : uint8_t {
    FK_SEL_LIT,      // SelectionKernelLiteral
    FK_SEL_COL,      // SelectionKernelColumns
    FK_LOGICAL,      // LogicalKernel
    FK_FILL,         // FillKernel
    FK_BINOP_CC,     // PerformOperationKernelColumns
    FK_BINOP_LC,     // PerformOperationKernelLiteralFirst
    FK_BINOP_CL,     // PerformOperationKernelLiteralSecond
    FK_PROBE_SEMI,   // FilterJoinKernel
    FK_PROBE_FULL,   // FullJoinKernel
    FK_BUILD_KEYS,   // BuildKeysHTKernel
    FK_BUILD_KEYVALS,// BuildKeyValsHTKernel
    FK_SUM,          // AggregateOperationKernel (terminal)
    FK_GROUPBY,      // GroupByAggregateKernel (terminal)
    // forms produced by the program optimiser (same per-row results):
    FK_SEL_RANGE,    // SEARCH over one range: logic(in, lit <= x <= lit2)
    FK_SEL_IN2,      // SEARCH over two points: logic(in, x == lit || x == lit2)
    FK_PROBE_BITS    // semi/full probe through a presence bitmap (+ value table)
};

// value table of FK_PROBE_BITS
//   FV_U16     16-bit offsets from vmin, one per slot (global memory); a slot
//              whose presence bit is clear holds VALS_UNSET, so the table
//              alone also answers the presence test
//   FV_PAIR    the reference (flag, value) pairs (global memory)
//   FV_RANK16  a rank dictionary staged in LDS at lds_off: the presence words
//              (u32 x ht_len words), the number of set bits before each word
//              (u16 x ht_len, padded to 4 bytes) and the values of the set
//              slots in slot order (u16 offsets from vmin); the value of a set
//              slot is values[prefix[w] + popcount(word & lower bits)]
enum FValKind : uint8_t { FV_NONE = 0, FV_U16 = 1, FV_PAIR = 2, FV_RANK16 = 3 };
constexpr uint32_t VALS_UNSET = 0xFFFFu;

// byte offsets of the parts of a rank dictionary of `words` presence words
__host__ __device__ constexpr uint32_t rank_prefix_off(uint32_t words) { return 4u * words; }
__host__ __device__ constexpr uint32_t rank_values_off(uint32_t words) { return 4u * words + (2u * words + 3u) / 4u * 4u; }

struct FInstr {
    uint8_t kind, cmp, logic, op;
    uint8_t fin, fout, floc, in_lds; // in_lds: FK_PROBE_BITS bitmap staged in LDS
    int8_t va, vb, vout, vkind;      // value registers (F_PF_BASE + j: prefetched
                                     // column j); -1 = memory operand a / b
    int32_t lit, lit2;
    float sel;                 // host estimate of the share of rows a predicate keeps
                               // (batched launches order their mask lookups by it)
    const int32_t *a;          // memory operand (row-indexed) when va < 0
    const int32_t *b;          // memory operand when vb < 0
    int32_t *store;            // write-back of vout (masked as the functor writes), or null
    const void *ht;            // reference-layout table; FK_PROBE_BITS: bitmap
    const void *vals;          // FK_PROBE_BITS value table (FV_U16: offsets from vmin; FV_PAIR: ht)
    int32_t ht_min, ht_max, ht_len, vmin; // FV_RANK16: ht_len = presence words
    uint32_t lds_off;          // byte offset of the staged bitmap / dictionary in dynamic LDS
    uint32_t lds_bytes;        // its size
};
static_assert(sizeof(FInstr) == 88, "FInstr: sel sits in the padding before the pointers");

enum FGroupMode : int { FG_GLOBAL = 0, FG_DENSE_LDS = 1, FG_HASH_LDS = 2 };

struct FGroup {
    int k;
    int lds;                   // FGroupMode
    int8_t key_reg[F_MAX_KEYS];
    int8_t agg_reg;
    int8_t pad[3];
    const int32_t *key_mem[F_MAX_KEYS];
    const int32_t *agg_mem;
    int32_t mins[F_MAX_KEYS];
    int32_t maxs[F_MAX_KEYS];
    uint32_t range[F_MAX_KEYS];
    int32_t *out_keys[F_MAX_KEYS];
    uint64_t *sums;
    uint32_t *seen;
    uint64_t prod;
    uint32_t lds_off;          // byte offset of the LDS aggregation area
    uint32_t hash_cap;         // FG_HASH_LDS: slots (power of two)
};

// Initial value of a flag register.  FF_UNDEF is 0 so that registers a
// program does not use (zero-initialised) are never loaded.
enum FFlagInit : uint8_t { FF_UNDEF = 0, FF_LOAD = 1, FF_TRUE = 2 };

constexpr uint32_t F_MAGIC = 0x5DB0F05Eu;

constexpr int F_MAX_BLOBS = 4;

struct FProgram {
    uint32_t magic;            // F_MAGIC; the kernel reads the program in place
                               // from its kernel-argument segment and checks it
    int ninstr;
    int vec;                   // 1: 4-row chunks with 16-byte column loads
    int nvals;                 // value registers used (LDS register file size)
    int nblobs;                // tables staged into LDS at kernel start
    uint32_t lds_bytes;        // dynamic LDS: registers, blobs, aggregation
    uint32_t vreg_off;         // byte offset of the register file
    uint32_t lds_base;         // batched launch: start of this program's LDS region
    int npf;                   // prefetched columns
    const int32_t *pf_mem[F_MAX_PF];
    const void *blob_src[F_MAX_BLOBS];
    uint32_t blob_bytes[F_MAX_BLOBS]; // multiples of 16
    uint32_t blob_off[F_MAX_BLOBS];
    uint8_t flag_init[F_MAX_FLAGS];
    const uint8_t *flag_load[F_MAX_FLAGS];
    uint8_t *flag_store[F_MAX_FLAGS];
    uint64_t *sum_acc;         // FK_SUM target
    FGroup gb;                 // FK_GROUPBY description
    FInstr ins[F_MAX_INSTR];
};

// LDS shared by the programs of one batched launch, and each program's share
// (a fixed share, so that a program's generated code does not depend on how
// many others it is batched with).
constexpr uint32_t F_BATCH_LDS = 128 * 1024;
constexpr uint32_t F_BATCH_MAX = 32;
constexpr uint32_t F_BATCH_LDS_EACH = F_BATCH_LDS / F_BATCH_MAX;

// Shared probe masks of a batched launch.  For one fact key column, a table
// over keys [lo, lo + len) of words whose bit i is 1 when every probe of
// program i on that column passes for the key (1 when program i has none);
// keys outside the table take `dflt` (only the bits of programs without a
// probe on the column).  A program's row can only reach its terminal where
// the AND of the row's words over all masked columns has its bit set.
enum FMaskKind : uint32_t {
    FM_BITS = 0, // presence bitmap, 32 slots per word
    FM_SEMI = 1, // the reference's bool table (BuildKeysHTKernel)
    FM_PAIR = 2  // the reference's (flag, value) pairs (BuildKeyValsHTKernel)
};
struct FMaskEntry {
    const void *ht;
    int32_t ht_min, ht_max;
    uint32_t bit, kind;
};
// A table stores, per key, only the window of `wbytes` bytes of the word at
// bit `shift` that covers the programs probing the column (programs take bit
// positions that keep those windows narrow); the bits outside the window are
// all ones (programs without a probe on the column).
struct FMaskCol {
    uint8_t *tab;          // len entries of wbytes bytes (padded to 4 bytes)
    int32_t lo;
    uint32_t len;
    uint32_t full;         // bits of every program of the launch
    uint32_t first, count; // its entries
    uint32_t shift, wbytes;
};
constexpr int F_MASK_MAX_COLS = 8;
constexpr int64_t F_MASK_MAX_LEN = int64_t(1) << 24;

// Rank dictionary of a batched mask table: the keys whose stored window is
// not zero (some probing program passes), in key order.  In global memory a
// 16-byte header (u32 ok: the entries fitted F_DICT_CAP; u32 count) precedes
// the part a block copies into LDS: the presence bitmap over the table's len
// keys (u32 words), the number of set bits before each word (u16, padded to
// 16 bytes) and the windows of the set keys (u32 each, up to F_DICT_CAP).
// A key whose presence bit is clear has window 0.
struct FMaskDict {
    const uint8_t *tab; // the mask table (len windows of wbytes bytes)
    uint8_t *dict;      // header, then the LDS part
    uint32_t len, wbytes;
};
constexpr uint32_t F_DICT_CAP = 4096;
constexpr uint32_t F_DICT_MAX_LEN = 4096 * 32; // bitmap words the builder holds in LDS
__host__ __device__ constexpr uint32_t dict_words(uint32_t len) { return (len + 31u) / 32u; }
__host__ __device__ constexpr uint32_t dict_prefix_off(uint32_t len) { return 4u * dict_words(len); }
__host__ __device__ constexpr uint32_t dict_vals_off(uint32_t len) { return (6u * dict_words(len) + 15u) / 16u * 16u; }
// bytes of the LDS part with room for F_DICT_CAP windows
__host__ __device__ constexpr uint32_t dict_lds_bytes(uint32_t len) { return dict_vals_off(len) + 4u * F_DICT_CAP; }

// Largest rank dictionary built for LDS staging.
constexpr uint32_t F_RANK_DICT_MAX = 32 * 1024;

// Dynamic LDS one fused block may use (registers, staged bitmaps, group-by).
constexpr uint32_t F_LDS_BUDGET = 96 * 1024;
// LDS aggregation sizes: dense slots up to F_DENSE_MAX, else a hash of
// F_HASH_CAP slots (power of two).
constexpr uint64_t F_DENSE_MAX = 1024;
constexpr uint32_t F_HASH_CAP = 1024;

} // namespace sdb
