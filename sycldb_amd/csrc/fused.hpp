// Fused execution of one KernelBundle (the per-segment operator sequence the
// reference queues in models/execution.hpp:500-651) as a single kernel.
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
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdb {

constexpr int F_MAX_INSTR = 24;
constexpr int F_MAX_FLAGS = 8;
constexpr int F_MAX_VALS = 4;
constexpr int F_MAX_KEYS = 4;

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
    FK_GROUPBY       // GroupByAggregateKernel (terminal)
};

struct FInstr {
    uint8_t kind, cmp, logic, op;
    uint8_t fin, fout, floc, pad0;
    int8_t va, vb, vout, pad1; // value registers; -1 = memory operand a / b
    int32_t lit;
    const int32_t *a;          // memory operand (row-indexed) when va < 0
    const int32_t *b;          // memory operand when vb < 0
    int32_t *store;            // write-back of vout (masked as the functor writes), or null
    const void *ht;
    int32_t ht_min, ht_max, ht_len, pad2;
};

struct FGroup {
    int k;
    int lds;                   // 1: LDS-privatised dense slots
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
};

// Initial value of a flag register.  FF_UNDEF is 0 so that registers a
// program does not use (zero-initialised) are never loaded.
enum FFlagInit : uint8_t { FF_UNDEF = 0, FF_LOAD = 1, FF_TRUE = 2 };

constexpr uint32_t F_MAGIC = 0x5DB0F05Eu;

struct FProgram {
    uint32_t magic;            // F_MAGIC; the kernel reads the program in place
                               // from its kernel-argument segment and checks it
    int ninstr;
    int vec;                   // 1: 4-row chunks with 16-byte column loads
    uint8_t flag_init[F_MAX_FLAGS];
    const uint8_t *flag_load[F_MAX_FLAGS];
    uint8_t *flag_store[F_MAX_FLAGS];
    uint64_t *sum_acc;         // FK_SUM target
    FGroup gb;                 // FK_GROUPBY description
    FInstr ins[F_MAX_INSTR];
};

// Launches the program over rows [0, n) of the segment the pointers address.
void launch_fused(const FProgram &p, int64_t n, hipStream_t s);
// Slots an LDS-privatised dense group-by can hold in the fused kernel.
uint64_t fused_groupby_lds_max_slots();

} // namespace sdb
