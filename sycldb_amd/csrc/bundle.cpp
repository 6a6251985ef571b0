// KernelBundle execution (models/execution.hpp:500-651): either one launch
// per functor (KernelData::execute, execution.hpp:260-497, with the bundle's
// kernels ordered on one stream instead of by waited events), or the whole
// bundle as one fused kernel.
#include <algorithm>
#include <map>

#include "engine.hpp"

namespace sdb {

namespace {

bool aligned(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

bool reads_virgin(const ExecContext &ctx, const uint8_t *p)
{
    return ctx.virgin_begin && !ctx.virgin_filled && p >= ctx.virgin_begin && p < ctx.virgin_end;
}

} // namespace

const char *kernel_type_name(KernelType t)
{
    switch (t) {
    case KernelType::EmptyKernel: return "EmptyKernel";
    case KernelType::LogicalKernel: return "LogicalKernel";
    case KernelType::SelectionKernelColumns: return "SelectionKernelColumns";
    case KernelType::SelectionKernelLiteral: return "SelectionKernelLiteral";
    case KernelType::FillKernel: return "FillKernel";
    case KernelType::PerformOperationKernelColumns: return "PerformOperationKernelColumns";
    case KernelType::PerformOperationKernelLiteralFirst: return "PerformOperationKernelLiteralFirst";
    case KernelType::PerformOperationKernelLiteralSecond: return "PerformOperationKernelLiteralSecond";
    case KernelType::BuildKeysHTKernel: return "BuildKeysHTKernel";
    case KernelType::FilterJoinKernel: return "FilterJoinKernel";
    case KernelType::BuildKeyValsHTKernel: return "BuildKeyValsHTKernel";
    case KernelType::FullJoinKernel: return "FullJoinKernel";
    case KernelType::AggregateOperationKernel: return "AggregateOperationKernel";
    case KernelType::GroupByAggregateKernel: return "GroupByAggregateKernel";
    }
    return "?";
}

void KernelBundle::execute(const ExecContext &ctx) const
{
    if (kernels_.empty()) return;
    Engine &eng = *ctx.eng;
    if (eng.fusion) {
        FProgram p;
        if (build_program(ctx, p)) {
            validate_program(p, kernels_.front().n, eng.registry);
            Engine::ProfRec rec;
            if (eng.profile) eng.prof_begin("fused:" + (ctx.table ? *ctx.table : std::string("?")), kernels_.front().n, rec);
            launch_fused(p, kernels_.front().n, ctx.stream);
            if (eng.profile) eng.prof_end(rec);
            eng.launches++;
            return;
        }
    }
    execute_unfused(ctx);
}

void KernelBundle::execute_unfused(const ExecContext &ctx) const
{
    hipStream_t s = ctx.stream;
    Engine &eng = *ctx.eng;
    // a flags buffer that is still logically all-true gets its values now
    bool need_fill = false;
    for (const KernelData &k : kernels_)
        need_fill |= reads_virgin(ctx, k.in_flags) || reads_virgin(ctx, k.local_flags);
    if (need_fill && !ctx.virgin_filled) {
        eng.registry.check(ctx.virgin_begin, (uint64_t)(ctx.virgin_end - ctx.virgin_begin), "initial flags");
        launch_fill_u8(const_cast<uint8_t *>(ctx.virgin_begin), 1, ctx.virgin_end - ctx.virgin_begin, s);
        eng.launches++;
        ctx.virgin_filled = true;
    }
    for (const KernelData &k : kernels_) {
        if (k.type == KernelType::EmptyKernel) continue;
        validate_kernel(k, eng.registry);
        Engine::ProfRec rec;
        if (eng.profile)
            eng.prof_begin(std::string("op:") + kernel_type_name(k.type) + ":" + (ctx.table ? *ctx.table : std::string("?")),
                           k.n, rec);
        switch (k.type) {
        case KernelType::EmptyKernel: continue;
        case KernelType::LogicalKernel: launch_logical(k.in_flags, k.local_flags, k.out_flags, k.n, k.logic, s); break;
        case KernelType::SelectionKernelColumns:
            launch_select_col(k.a, k.b, k.in_flags, k.out_flags, k.n, k.cmp, k.logic, s);
            break;
        case KernelType::SelectionKernelLiteral:
            launch_select_lit(k.a, k.in_flags, k.out_flags, k.n, k.cmp, k.logic, k.literal, s);
            break;
        case KernelType::FillKernel: launch_fill_i32(k.result, k.literal, k.n, s); break;
        case KernelType::PerformOperationKernelColumns:
            launch_binop(0, k.op, k.result, k.a, k.b, 0, k.in_flags, k.n, s);
            break;
        case KernelType::PerformOperationKernelLiteralFirst:
            launch_binop(1, k.op, k.result, nullptr, k.b, k.literal, k.in_flags, k.n, s);
            break;
        case KernelType::PerformOperationKernelLiteralSecond:
            launch_binop(2, k.op, k.result, k.a, nullptr, k.literal, k.in_flags, k.n, s);
            break;
        case KernelType::BuildKeysHTKernel:
            launch_build_keys(k.a, k.in_flags, k.n, static_cast<uint8_t *>(k.ht), k.ht_len, k.ht_min, s);
            break;
        case KernelType::FilterJoinKernel:
            launch_probe_semi(k.a, k.in_flags, k.out_flags, k.n, static_cast<const uint8_t *>(k.ht), k.ht_min,
                              k.ht_max, s);
            break;
        case KernelType::BuildKeyValsHTKernel:
            launch_build_keyvals(k.a, k.b, k.in_flags, k.n, static_cast<int32_t *>(k.ht), k.ht_len, k.ht_min, s);
            break;
        case KernelType::FullJoinKernel:
            launch_probe_full(k.a, k.result, k.in_flags, k.out_flags, k.n, static_cast<const int32_t *>(k.ht),
                              k.ht_min, k.ht_max, s);
            break;
        case KernelType::AggregateOperationKernel: launch_sum_flagged(k.a, k.in_flags, k.n, k.agg_res, s); break;
        case KernelType::GroupByAggregateKernel: {
            const GroupBySpec &g = *k.gb;
            launch_groupby_dense(g.keys.data(), (int)g.keys.size(), g.mins.data(), g.maxs.data(), k.a, k.in_flags,
                                 k.n, g.out_keys.data(), g.sums, g.seen, g.prod, s, 0);
            break;
        }
        }
        if (eng.profile) eng.prof_end(rec);
        eng.launches++;
    }
}

// Translates the bundle into a fused program.  Returns false when the bundle
// uses something the fused kernel does not cover (then it runs unfused).
bool KernelBundle::build_program(const ExecContext &ctx, FProgram &p) const
{
    p = FProgram{};
    p.magic = F_MAGIC;
    const int64_t n = kernels_.front().n;
    for (const KernelData &k : kernels_)
        if (k.n != n) return false;

    std::vector<const uint8_t *> fptr;     // flag register -> buffer
    std::vector<bool> fwritten;
    std::map<const int32_t *, int> vmap;   // column produced in this bundle -> value register
    bool vec = true;
    bool terminal_seen = false;
    bool ok = true;

    auto freg = [&](const uint8_t *ptr, bool write) -> int {
        for (size_t j = 0; j < fptr.size(); j++)
            if (fptr[j] == ptr) {
                if (write) fwritten[j] = true;
                return (int)j;
            }
        if ((int)fptr.size() >= F_MAX_FLAGS) {
            ok = false;
            return 0;
        }
        const int j = (int)fptr.size();
        fptr.push_back(ptr);
        fwritten.push_back(write);
        if (write) {
            p.flag_init[j] = FF_UNDEF;
        } else if (reads_virgin(ctx, ptr)) {
            p.flag_init[j] = FF_TRUE;
        } else {
            p.flag_init[j] = FF_LOAD;
            p.flag_load[j] = ptr;
            vec &= aligned(ptr, 4);
        }
        return j;
    };
    // operand: value register when produced earlier in the bundle, else memory
    auto src = [&](const int32_t *ptr, int8_t &reg, const int32_t *&mem) {
        auto it = vmap.find(ptr);
        if (it != vmap.end()) {
            reg = (int8_t)it->second;
            mem = nullptr;
        } else {
            reg = -1;
            mem = ptr;
            vec &= aligned(ptr, 16);
        }
    };
    auto dst = [&](int32_t *ptr, FInstr &I) {
        auto it = vmap.find(ptr);
        int r;
        if (it != vmap.end()) {
            r = it->second;
        } else {
            r = (int)vmap.size();
            if (r >= F_MAX_VALS) {
                ok = false;
                r = 0;
            }
            vmap[ptr] = r;
        }
        I.vout = (int8_t)r;
        I.store = ctx.live->contains(ptr) ? ptr : nullptr;
        if (I.store) vec &= aligned(ptr, 16);
    };

    int ni = 0;
    for (const KernelData &k : kernels_) {
        if (k.type == KernelType::EmptyKernel) continue;
        if (ni >= F_MAX_INSTR || terminal_seen) return false;
        FInstr &I = p.ins[ni];
        I = FInstr{};
        I.va = I.vb = I.vout = -1;
        switch (k.type) {
        case KernelType::SelectionKernelLiteral:
        case KernelType::SelectionKernelColumns:
            I.kind = k.type == KernelType::SelectionKernelLiteral ? FK_SEL_LIT : FK_SEL_COL;
            I.cmp = (uint8_t)k.cmp;
            I.logic = (uint8_t)k.logic;
            I.lit = k.literal;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            if (I.kind == FK_SEL_COL) src(k.b, I.vb, I.b);
            I.fout = (uint8_t)freg(k.out_flags, true);
            break;
        case KernelType::LogicalKernel:
            I.kind = FK_LOGICAL;
            I.logic = (uint8_t)k.logic;
            I.fin = (uint8_t)freg(k.in_flags, false);
            I.floc = (uint8_t)freg(k.local_flags, false);
            I.fout = (uint8_t)freg(k.out_flags, true);
            break;
        case KernelType::FillKernel:
            I.kind = FK_FILL;
            I.lit = k.literal;
            dst(k.result, I);
            break;
        case KernelType::PerformOperationKernelColumns:
        case KernelType::PerformOperationKernelLiteralFirst:
        case KernelType::PerformOperationKernelLiteralSecond:
            I.kind = k.type == KernelType::PerformOperationKernelColumns        ? FK_BINOP_CC
                     : k.type == KernelType::PerformOperationKernelLiteralFirst ? FK_BINOP_LC
                                                                                 : FK_BINOP_CL;
            I.op = (uint8_t)k.op;
            I.lit = k.literal;
            I.fin = (uint8_t)freg(k.in_flags, false);
            if (I.kind != FK_BINOP_LC) src(k.a, I.va, I.a);
            if (I.kind != FK_BINOP_CL) src(k.b, I.vb, I.b);
            dst(k.result, I);
            break;
        case KernelType::FilterJoinKernel:
            I.kind = FK_PROBE_SEMI;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_max = k.ht_max;
            I.fout = (uint8_t)freg(k.out_flags, true);
            break;
        case KernelType::FullJoinKernel:
            I.kind = FK_PROBE_FULL;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_max = k.ht_max;
            if (!aligned(k.ht, 8)) return false;
            dst(k.result, I);
            I.fout = (uint8_t)freg(k.out_flags, true);
            break;
        case KernelType::BuildKeysHTKernel:
            I.kind = FK_BUILD_KEYS;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_len = k.ht_len;
            break;
        case KernelType::BuildKeyValsHTKernel:
            I.kind = FK_BUILD_KEYVALS;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            src(k.b, I.vb, I.b);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_len = k.ht_len;
            if (!aligned(k.ht, 8)) return false;
            break;
        case KernelType::AggregateOperationKernel:
            I.kind = FK_SUM;
            I.fin = (uint8_t)freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            p.sum_acc = k.agg_res;
            if (!aligned(k.agg_res, 8)) return false;
            terminal_seen = true;
            break;
        case KernelType::GroupByAggregateKernel: {
            const GroupBySpec &g = *k.gb;
            const int nk = (int)g.keys.size();
            if (nk < 1 || nk > F_MAX_KEYS) return false;
            I.kind = FK_GROUPBY;
            I.fin = (uint8_t)freg(k.in_flags, false);
            FGroup &G = p.gb;
            G.k = nk;
            uint64_t exact = 1;
            bool fits = true;
            for (int j = 0; j < nk; j++) {
                src(g.keys[j], G.key_reg[j], G.key_mem[j]);
                G.mins[j] = g.mins[j];
                G.maxs[j] = g.maxs[j];
                G.range[j] = (uint32_t)g.maxs[j] - (uint32_t)g.mins[j] + 1u;
                G.out_keys[j] = g.out_keys[j];
                if (g.maxs[j] < g.mins[j]) fits = false;
                exact *= (uint64_t)((int64_t)g.maxs[j] - (int64_t)g.mins[j] + 1);
                if (exact > g.prod) fits = false;
            }
            src(k.a, G.agg_reg, G.agg_mem);
            G.sums = g.sums;
            G.seen = g.seen;
            G.prod = g.prod;
            if (g.prod == 0) return false;
            G.lds = fits && exact == g.prod && g.prod <= fused_groupby_lds_max_slots();
            if (!aligned(g.sums, 8)) return false;
            terminal_seen = true;
            break;
        }
        default: return false;
        }
        if (!ok) return false;
        ni++;
    }
    if (ni == 0) return false;
    p.ninstr = ni;
    // write back every flag buffer the bundle changed (or still logically
    // all-true) that is referenced afterwards
    for (size_t j = 0; j < fptr.size(); j++) {
        const bool changed = fwritten[j] || p.flag_init[j] == FF_TRUE;
        if (changed && ctx.live->contains(fptr[j])) {
            p.flag_store[j] = const_cast<uint8_t *>(fptr[j]);
            vec &= aligned(fptr[j], 4);
        }
    }
    p.vec = vec ? 1 : 0;
    return true;
}

} // namespace sdb
