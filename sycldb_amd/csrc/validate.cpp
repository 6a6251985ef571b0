// Host-side bounds checks run before every engine launch: each buffer a
// kernel will touch must lie inside one registered device allocation and be
// long enough for the rows (or hash-table slots, group slots) it is indexed
// with.  A violation throws instead of launching.
#include <string>

#include "engine.hpp"

namespace sdb {

void AllocRegistry::add(const void *p, uint64_t bytes)
{
    if (p) allocs_[(uintptr_t)p] = bytes;
}

void AllocRegistry::remove(const void *p) { allocs_.erase((uintptr_t)p); }

void AllocRegistry::check(const void *p, uint64_t bytes, const char *what) const
{
    if (bytes == 0) return;
    if (!p) fail(SDB_ERR_RUNTIME, std::string("launch check: null ") + what);
    const uintptr_t a = (uintptr_t)p;
    auto it = allocs_.upper_bound(a);
    if (it == allocs_.begin())
        fail(SDB_ERR_RUNTIME, std::string("launch check: ") + what + " is not in a device allocation");
    --it;
    if (a < it->first || a + bytes > it->first + it->second || a + bytes < a)
        fail(SDB_ERR_RUNTIME, std::string("launch check: ") + what + " (" + std::to_string(bytes) +
                                  " bytes) overruns its device allocation");
}

namespace {

uint64_t range_len(int32_t mn, int32_t mx)
{
    if (mx < mn) fail(SDB_ERR_RUNTIME, "launch check: hash table max < min");
    return (uint64_t)((int64_t)mx - (int64_t)mn + 1);
}

} // namespace

void validate_kernel(const KernelData &k, const AllocRegistry &reg)
{
    const uint64_t n = (uint64_t)k.n;
    if (k.n < 0) fail(SDB_ERR_RUNTIME, "launch check: negative row count");
    if (n == 0) return;
    const uint64_t col = 4 * n;
    switch (k.type) {
    case KernelType::EmptyKernel: return;
    case KernelType::LogicalKernel:
        reg.check(k.in_flags, n, "input flags");
        reg.check(k.local_flags, n, "local flags");
        reg.check(k.out_flags, n, "output flags");
        return;
    case KernelType::SelectionKernelColumns: reg.check(k.b, col, "second column"); [[fallthrough]];
    case KernelType::SelectionKernelLiteral:
        reg.check(k.a, col, "column");
        reg.check(k.in_flags, n, "input flags");
        reg.check(k.out_flags, n, "output flags");
        return;
    case KernelType::FillKernel: reg.check(k.result, col, "fill target"); return;
    case KernelType::PerformOperationKernelColumns:
    case KernelType::PerformOperationKernelLiteralFirst:
    case KernelType::PerformOperationKernelLiteralSecond:
        if (k.type != KernelType::PerformOperationKernelLiteralFirst) reg.check(k.a, col, "first operand");
        if (k.type != KernelType::PerformOperationKernelLiteralSecond) reg.check(k.b, col, "second operand");
        reg.check(k.result, col, "result column");
        reg.check(k.in_flags, n, "flags");
        return;
    case KernelType::BuildKeysHTKernel:
        if (k.ht_len <= 0) fail(SDB_ERR_RUNTIME, "launch check: hash table length");
        reg.check(k.a, col, "build keys");
        reg.check(k.in_flags, n, "build flags");
        reg.check(k.ht, (uint64_t)k.ht_len, "key hash table");
        return;
    case KernelType::BuildKeyValsHTKernel:
        if (k.ht_len <= 0) fail(SDB_ERR_RUNTIME, "launch check: hash table length");
        reg.check(k.a, col, "build keys");
        reg.check(k.b, col, "build values");
        reg.check(k.in_flags, n, "build flags");
        reg.check(k.ht, 8 * (uint64_t)k.ht_len, "key-value hash table");
        return;
    case KernelType::FilterJoinKernel:
        reg.check(k.a, col, "probe keys");
        reg.check(k.in_flags, n, "input flags");
        reg.check(k.out_flags, n, "output flags");
        reg.check(k.ht, range_len(k.ht_min, k.ht_max), "key hash table");
        return;
    case KernelType::FullJoinKernel:
        reg.check(k.a, col, "probe keys");
        reg.check(k.result, col, "join value column");
        reg.check(k.in_flags, n, "input flags");
        reg.check(k.out_flags, n, "output flags");
        reg.check(k.ht, 8 * range_len(k.ht_min, k.ht_max), "key-value hash table");
        return;
    case KernelType::AggregateOperationKernel:
        reg.check(k.a, col, "aggregated column");
        reg.check(k.in_flags, n, "flags");
        reg.check(k.agg_res, 8, "accumulator");
        return;
    case KernelType::GroupByAggregateKernel: {
        if (!k.gb) fail(SDB_ERR_RUNTIME, "launch check: group-by without spec");
        const GroupBySpec &g = *k.gb;
        if (g.prod == 0 || g.keys.empty() || g.keys.size() != g.out_keys.size())
            fail(SDB_ERR_RUNTIME, "launch check: malformed group-by");
        reg.check(k.a, col, "aggregated column");
        reg.check(k.in_flags, n, "flags");
        for (size_t j = 0; j < g.keys.size(); j++) {
            reg.check(g.keys[j], col, "group key column");
            reg.check(g.out_keys[j], 4 * g.prod, "group key output");
        }
        reg.check(g.sums, 8 * g.prod, "group sums");
        reg.check(g.seen, 4 * g.prod, "group seen flags");
        return;
    }
    }
}

void validate_program(const FProgram &p, int64_t n_rows, const AllocRegistry &reg)
{
    if (p.magic != F_MAGIC) fail(SDB_ERR_RUNTIME, "launch check: program magic");
    if (p.ninstr < 1 || p.ninstr > F_MAX_INSTR) fail(SDB_ERR_RUNTIME, "launch check: instruction count");
    if (p.lds_bytes > F_LDS_BUDGET) fail(SDB_ERR_RUNTIME, "launch check: LDS budget");
    if (p.nvals < 0 || p.nvals > F_MAX_VALS) fail(SDB_ERR_RUNTIME, "launch check: register count");
    auto lds_area = [&](uint64_t off, uint64_t bytes, const char *what) {
        if (off + bytes > p.lds_bytes) fail(SDB_ERR_RUNTIME, std::string("launch check: LDS area of ") + what);
    };
    lds_area(p.vreg_off, (uint64_t)p.nvals * 256 * 16, "registers");
    if (p.nblobs < 0 || p.nblobs > F_MAX_BLOBS) fail(SDB_ERR_RUNTIME, "launch check: blob count");
    for (int b = 0; b < p.nblobs; b++) {
        if (p.blob_bytes[b] % 16) fail(SDB_ERR_RUNTIME, "launch check: blob size");
        reg.check(p.blob_src[b], p.blob_bytes[b], "staged table");
        lds_area(p.blob_off[b], p.blob_bytes[b], "a staged table");
    }
    if (n_rows <= 0) return;
    const uint64_t n = (uint64_t)n_rows, col = 4 * n;
    for (int j = 0; j < F_MAX_FLAGS; j++) {
        if (p.flag_init[j] > FF_TRUE) fail(SDB_ERR_RUNTIME, "launch check: flag init");
        if (p.flag_init[j] == FF_LOAD) reg.check(p.flag_load[j], n, "loaded flags");
        if (p.flag_store[j]) reg.check(p.flag_store[j], n, "stored flags");
    }
    auto fl = [&](uint8_t r) {
        if (r >= F_MAX_FLAGS) fail(SDB_ERR_RUNTIME, "launch check: flag register");
    };
    auto vreg = [&](int8_t r) {
        if (r < 0 || r >= p.nvals) fail(SDB_ERR_RUNTIME, "launch check: value register");
    };
    if (p.npf < 0 || p.npf > F_MAX_PF) fail(SDB_ERR_RUNTIME, "launch check: prefetch count");
    for (int j = 0; j < p.npf; j++) {
        reg.check(p.pf_mem[j], col, "prefetched column");
        if (p.vec && (uintptr_t)p.pf_mem[j] % 16) fail(SDB_ERR_RUNTIME, "launch check: prefetched column alignment");
    }
    auto operand = [&](int8_t r, const int32_t *mem, const char *what) {
        if (r >= F_PF_BASE) {
            if (r - F_PF_BASE >= p.npf) fail(SDB_ERR_RUNTIME, "launch check: prefetch register");
        } else if (r >= 0) {
            vreg(r);
        } else {
            reg.check(mem, col, what);
        }
    };
    bool has_sum = false, has_gb = false;
    for (int i = 0; i < p.ninstr; i++) {
        const FInstr &I = p.ins[i];
        fl(I.fin);
        switch (I.kind) {
        case FK_SEL_LIT:
        case FK_SEL_COL:
        case FK_SEL_RANGE:
        case FK_SEL_IN2:
            if (I.cmp > SDB_GE || I.logic > SDB_OR) fail(SDB_ERR_RUNTIME, "launch check: selection op");
            operand(I.va, I.a, "selection column");
            if (I.kind == FK_SEL_COL) operand(I.vb, I.b, "selection column");
            fl(I.fout);
            break;
        case FK_LOGICAL:
            if (I.logic > SDB_OR) fail(SDB_ERR_RUNTIME, "launch check: logical op");
            fl(I.floc);
            fl(I.fout);
            break;
        case FK_FILL:
            vreg(I.vout);
            if (I.store) reg.check(I.store, col, "fill store");
            break;
        case FK_BINOP_CC:
        case FK_BINOP_LC:
        case FK_BINOP_CL:
            if (I.op > SDB_SUB) fail(SDB_ERR_RUNTIME, "launch check: arithmetic op");
            if (I.kind != FK_BINOP_LC) operand(I.va, I.a, "arithmetic operand");
            if (I.kind != FK_BINOP_CL) operand(I.vb, I.b, "arithmetic operand");
            vreg(I.vout);
            if (I.store) reg.check(I.store, col, "arithmetic store");
            break;
        case FK_PROBE_SEMI:
            operand(I.va, I.a, "probe keys");
            reg.check(I.ht, range_len(I.ht_min, I.ht_max), "semi-join table");
            fl(I.fout);
            break;
        case FK_PROBE_FULL:
            operand(I.va, I.a, "probe keys");
            reg.check(I.ht, 8 * range_len(I.ht_min, I.ht_max), "join table");
            vreg(I.vout);
            if (I.store) reg.check(I.store, col, "join value store");
            fl(I.fout);
            break;
        case FK_PROBE_BITS: {
            operand(I.va, I.a, "probe keys");
            const uint64_t len = range_len(I.ht_min, I.ht_max);
            const uint64_t bytes = (len + 31) / 32 * 4;
            reg.check(I.ht, bytes, "probe bitmap");
            if (I.in_lds) lds_area(I.lds_off, bytes, "a probe bitmap");
            if (I.vkind == FV_U16) {
                reg.check(I.vals, 2 * len, "probe value table");
            } else if (I.vkind == FV_PAIR) {
                reg.check(I.vals, 8 * len, "probe value pairs");
            } else if (I.vkind == FV_RANK16) {
                // the dictionary: words, prefix counts and at least one value
                // per set slot lie in its staged LDS area
                if (!I.in_lds || (uint64_t)I.ht_len != (len + 31) / 32)
                    fail(SDB_ERR_RUNTIME, "launch check: rank dictionary shape");
                if (I.lds_bytes < rank_values_off((uint32_t)I.ht_len))
                    fail(SDB_ERR_RUNTIME, "launch check: rank dictionary size");
                lds_area(I.lds_off, I.lds_bytes, "a rank dictionary");
            } else if (I.vkind != FV_NONE) {
                fail(SDB_ERR_RUNTIME, "launch check: value table kind");
            }
            if (I.vkind != FV_NONE) vreg(I.vout);
            if (I.store) reg.check(I.store, col, "join value store");
            fl(I.fout);
            break;
        }
        case FK_BUILD_KEYS:
            operand(I.va, I.a, "build keys");
            if (I.ht_len <= 0) fail(SDB_ERR_RUNTIME, "launch check: table length");
            reg.check(I.ht, (uint64_t)I.ht_len, "key table");
            break;
        case FK_BUILD_KEYVALS:
            operand(I.va, I.a, "build keys");
            operand(I.vb, I.b, "build values");
            if (I.ht_len <= 0) fail(SDB_ERR_RUNTIME, "launch check: table length");
            reg.check(I.ht, 8 * (uint64_t)I.ht_len, "key-value table");
            break;
        case FK_SUM:
            operand(I.va, I.a, "summed column");
            reg.check(p.sum_acc, 8, "sum accumulator");
            has_sum = true;
            break;
        case FK_GROUPBY: {
            const FGroup &g = p.gb;
            if (g.k < 1 || g.k > F_MAX_KEYS || g.prod == 0) fail(SDB_ERR_RUNTIME, "launch check: group-by shape");
            for (int j = 0; j < g.k; j++) {
                operand(g.key_reg[j], g.key_mem[j], "group key");
                reg.check(g.out_keys[j], 4 * g.prod, "group key output");
            }
            operand(g.agg_reg, g.agg_mem, "aggregated column");
            reg.check(g.sums, 8 * g.prod, "group sums");
            reg.check(g.seen, 4 * g.prod, "group seen flags");
            if (g.lds == FG_DENSE_LDS) {
                if (g.prod > F_DENSE_MAX) fail(SDB_ERR_RUNTIME, "launch check: dense LDS slots");
                if (g.k > F_MAX_KEYS) fail(SDB_ERR_RUNTIME, "launch check: dense LDS keys");
                lds_area(g.lds_off, g.prod * 8 + (g.prod + 31) / 32 * 4, "the dense aggregation");
            } else if (g.lds == FG_HASH_LDS) {
                if (g.hash_cap == 0 || (g.hash_cap & (g.hash_cap - 1)) || g.prod > 0x80000000ull)
                    fail(SDB_ERR_RUNTIME, "launch check: hash aggregation shape");
                lds_area(g.lds_off, (uint64_t)g.hash_cap * 12, "the hash aggregation");
            } else if (g.lds != FG_GLOBAL) {
                fail(SDB_ERR_RUNTIME, "launch check: aggregation mode");
            }
            has_gb = true;
            break;
        }
        default: fail(SDB_ERR_RUNTIME, "launch check: unknown instruction");
        }
    }
    if (p.sum_acc && !has_sum) fail(SDB_ERR_RUNTIME, "launch check: accumulator without sum");
    if (p.gb.prod && !has_gb) fail(SDB_ERR_RUNTIME, "launch check: group spec without group-by");
}

} // namespace sdb
