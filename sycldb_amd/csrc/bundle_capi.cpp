// sdb_bundle_exec: one KernelBundle handed over by the caller as a list of
// operators, run fused.
//
// The reference's KernelBundle::execute (models/execution.hpp:528-650) has
// a fused path, switched off there (disable_fusion, 546): it packs a prefix
// of at most eight fusable kernels into the slots of a SYCLDBContext
// (kernels/jit_kernels.hpp:12-20; params[i*8+k]: comparison, logic,
// arithmetic op, min, max, length) and threads one `pass` flag through their
// row functions (kernels/jit_kernels_impl.hpp:23-90).  Here the caller
// describes every kernel of the bundle with the arguments of its KernelData
// (sdb_op) and the whole bundle runs through this library's program builder
// as one kernel: every flag buffer and produced column becomes a register,
// and every buffer an operator writes is written back, so the result is the
// unfused sequence's.  A bundle the fused kernel does not cover runs one
// kernel per operator, in order, on the same stream.
//
// The caller owns every buffer.  Their sizes are the ones the operators
// imply (n rows for columns and flags, ht_len slots for tables, prod slots
// for group-by arrays), and every launch is checked against them.
#include <cstdio>
#include <memory>
#include <string>

#include "capi_util.hpp"
#include "engine.hpp"

using namespace sdb;

namespace {

KernelType kernel_type_of(int kind)
{
    switch (kind) {
    case SDB_OP_EMPTY: return KernelType::EmptyKernel;
    case SDB_OP_LOGICAL: return KernelType::LogicalKernel;
    case SDB_OP_SELECT_COL: return KernelType::SelectionKernelColumns;
    case SDB_OP_SELECT_LIT: return KernelType::SelectionKernelLiteral;
    case SDB_OP_FILL: return KernelType::FillKernel;
    case SDB_OP_BINOP_CC: return KernelType::PerformOperationKernelColumns;
    case SDB_OP_BINOP_LC: return KernelType::PerformOperationKernelLiteralFirst;
    case SDB_OP_BINOP_CL: return KernelType::PerformOperationKernelLiteralSecond;
    case SDB_OP_BUILD_KEYS: return KernelType::BuildKeysHTKernel;
    case SDB_OP_PROBE_SEMI: return KernelType::FilterJoinKernel;
    case SDB_OP_BUILD_KEYVALS: return KernelType::BuildKeyValsHTKernel;
    case SDB_OP_PROBE_FULL: return KernelType::FullJoinKernel;
    case SDB_OP_SUM: return KernelType::AggregateOperationKernel;
    case SDB_OP_GROUPBY: return KernelType::GroupByAggregateKernel;
    default: fail(SDB_ERR_INVALID_ARG, "sdb_bundle_exec: unknown operator kind " + std::to_string(kind));
    }
}

// The operator as the engine's KernelData, its buffers registered with the
// sizes it implies, and the buffers it writes marked live.
KernelData kernel_of(const sdb_op &o, int64_t n, AllocRegistry &reg, LiveSet &live)
{
    KernelData k;
    k.type = kernel_type_of(o.kind);
    k.n = n;
    const uint64_t rows = (uint64_t)n, col = 4 * rows;
    auto in = [&](const void *p, uint64_t bytes, const char *what) {
        if (!p && bytes) fail(SDB_ERR_INVALID_ARG, std::string("sdb_bundle_exec: null ") + what);
        reg.add(p, bytes);
    };
    auto out = [&](const void *p, uint64_t bytes, const char *what) {
        in(p, bytes, what);
        live.add(p, bytes);
    };
    auto cmp_logic = [&] {
        require(o.cmp >= SDB_EQ && o.cmp <= SDB_GE, "sdb_bundle_exec: comparison op out of range");
        require(o.logic >= SDB_NONE && o.logic <= SDB_OR, "sdb_bundle_exec: logical op out of range");
        k.cmp = o.cmp;
        k.logic = o.logic;
    };
    auto binop = [&] {
        require(o.binop >= SDB_MUL && o.binop <= SDB_SUB, "sdb_bundle_exec: arithmetic op out of range");
        k.op = o.binop;
    };
    auto table = [&](bool pairs, bool from_bounds) {
        int64_t len = o.ht_len;
        if (from_bounds && len == 0) len = (int64_t)o.ht_max - o.ht_min + 1;
        require(len > 0 && len <= INT32_MAX, "sdb_bundle_exec: hash table length out of range");
        k.ht_len = (int32_t)len;
        k.ht_min = o.ht_min;
        k.ht_max = from_bounds ? o.ht_max : (int32_t)((int64_t)o.ht_min + len - 1);
        k.ht = o.ht;
    };
    switch (k.type) {
    case KernelType::EmptyKernel: break;
    case KernelType::LogicalKernel:
        require(o.logic >= SDB_NONE && o.logic <= SDB_OR, "sdb_bundle_exec: logical op out of range");
        k.logic = o.logic;
        k.in_flags = o.in_flags, in(o.in_flags, rows, "input flags");
        k.local_flags = o.local_flags, in(o.local_flags, rows, "local flags");
        k.out_flags = o.out_flags, out(o.out_flags, rows, "output flags");
        break;
    case KernelType::SelectionKernelColumns:
        k.b = o.col2, in(o.col2, col, "second column");
        [[fallthrough]];
    case KernelType::SelectionKernelLiteral:
        cmp_logic();
        k.literal = o.value;
        k.a = o.col, in(o.col, col, "column");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "input flags");
        k.out_flags = o.out_flags, out(o.out_flags, rows, "output flags");
        break;
    case KernelType::FillKernel:
        k.literal = o.value;
        k.result = o.res, out(o.res, col, "fill target");
        break;
    case KernelType::PerformOperationKernelColumns:
    case KernelType::PerformOperationKernelLiteralFirst:
    case KernelType::PerformOperationKernelLiteralSecond:
        binop();
        k.literal = o.value;
        if (k.type == KernelType::PerformOperationKernelColumns) {
            k.a = o.col, in(o.col, col, "first operand");
            k.b = o.col2, in(o.col2, col, "second operand");
        } else if (k.type == KernelType::PerformOperationKernelLiteralFirst) {
            k.b = o.col, in(o.col, col, "column operand"); // value op col
        } else {
            k.a = o.col, in(o.col, col, "column operand"); // col op value
        }
        k.in_flags = o.in_flags, in(o.in_flags, rows, "flags");
        k.result = o.res, out(o.res, col, "result column");
        break;
    case KernelType::BuildKeysHTKernel:
        table(false, false);
        k.a = o.col, in(o.col, col, "build keys");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "build flags");
        out(o.ht, (uint64_t)k.ht_len, "key hash table");
        break;
    case KernelType::BuildKeyValsHTKernel:
        table(true, false);
        k.a = o.col, in(o.col, col, "build keys");
        k.b = o.col2, in(o.col2, col, "build values");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "build flags");
        out(o.ht, 8 * (uint64_t)k.ht_len, "key-value hash table");
        break;
    case KernelType::FilterJoinKernel:
        table(false, true);
        k.a = o.col, in(o.col, col, "probe column");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "input flags");
        k.out_flags = o.out_flags, out(o.out_flags, rows, "output flags");
        in(o.ht, (uint64_t)k.ht_len, "key hash table");
        break;
    case KernelType::FullJoinKernel:
        table(true, true);
        k.a = o.col, in(o.col, col, "probe column");
        k.result = o.res, out(o.res, col, "joined value column");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "input flags");
        k.out_flags = o.out_flags, out(o.out_flags, rows, "output flags");
        in(o.ht, 8 * (uint64_t)k.ht_len, "key-value hash table");
        break;
    case KernelType::AggregateOperationKernel:
        k.a = o.col, in(o.col, col, "summed column");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "flags");
        k.agg_res = o.acc, out(o.acc, 8, "accumulator");
        break;
    case KernelType::GroupByAggregateKernel: {
        require(o.nkeys >= 1 && o.nkeys <= SDB_OP_MAX_KEYS, "sdb_bundle_exec: group-by key count out of range");
        require(o.prod >= 1 && o.prod <= (1ull << 40), "sdb_bundle_exec: group-by slot count out of range");
        auto g = std::make_shared<GroupBySpec>();
        for (int j = 0; j < o.nkeys; j++) {
            in(o.keys[j], col, "group-by key column");
            out(o.out_keys[j], 4 * o.prod, "group-by key output");
            g->keys.push_back(o.keys[j]);
            g->mins.push_back(o.mins[j]);
            g->maxs.push_back(o.maxs[j]);
            g->out_keys.push_back(o.out_keys[j]);
        }
        g->sums = o.sums, out(o.sums, 8 * o.prod, "group-by sums");
        g->seen = o.seen, out(o.seen, 4 * o.prod, "group-by seen flags");
        g->prod = o.prod;
        k.gb = g;
        k.a = o.col, in(o.col, col, "aggregated column");
        k.in_flags = o.in_flags, in(o.in_flags, rows, "flags");
        break;
    }
    }
    return k;
}

} // namespace

extern "C" {

int sdb_bundle_exec(const sdb_op *ops, int nops, int64_t n, sdb_stream stream)
{
    return guarded([&] {
        require(nops >= 0 && nops <= SDB_BUNDLE_MAX_OPS, "sdb_bundle_exec: operator count out of range");
        require(nops == 0 || ops, "sdb_bundle_exec: null operator list");
        require(n >= 0 && n <= INT32_MAX, "sdb_bundle_exec: row count out of range (at most 2^31-1 per bundle)");
        if (nops == 0 || n == 0) return;
        // an engine without tables or allocator: its registry bounds the
        // caller's buffers, its settings are the defaults (fusion on, the
        // autotuned compiled kernels, the interpreter meanwhile)
        Engine eng;
        LiveSet live;
        KernelBundle b;
        for (int i = 0; i < nops; i++) b.add_kernel(kernel_of(ops[i], n, eng.registry, live));
        SDB_HIP(hipGetDevice(&eng.device));
        eng.stream = static_cast<hipStream_t>(stream);
        static const std::string table = "bundle";
        ExecContext ctx{&eng, static_cast<hipStream_t>(stream), &live, &table};
        b.execute(ctx);
    });
}

} // extern "C"
