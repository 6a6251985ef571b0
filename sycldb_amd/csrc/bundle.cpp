// KernelBundle execution (models/execution.hpp:500-651): either one launch
// per functor (KernelData::execute, execution.hpp:260-497, with the bundle's
// kernels ordered on one stream instead of by waited events), or the whole
// bundle as one fused kernel.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cmath>
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

static const char *fkind_name(uint8_t k)
{
    static const char *names[] = {"SEL_LIT",  "SEL_COL",    "LOGICAL", "FILL",       "BINOP_CC",
                                  "BINOP_LC", "BINOP_CL",   "PROBE_SEMI", "PROBE_FULL", "BUILD_KEYS",
                                  "BUILD_KEYVALS", "SUM",   "GROUPBY", "SEL_RANGE",  "SEL_IN2", "PROBE_BITS"};
    return k < sizeof(names) / sizeof(names[0]) ? names[k] : "?";
}

// SDB_DEBUG_FUSED=1 prints each fused program to stderr.
static void dump_program(const FProgram &p, int64_t n, const std::string &table)
{
    static const bool on = [] {
        const char *e = getenv("SDB_DEBUG_FUSED");
        return e && *e && *e != '0';
    }();
    if (!on) return;
    fprintf(stderr, "[fused %s] rows=%lld vec=%d instr=%d vals=%d prefetch=%d blobs=%d lds=%u group=%d\n",
            table.c_str(), (long long)n, p.vec, p.ninstr, p.nvals, p.npf, p.nblobs, p.lds_bytes, p.gb.lds);
    for (int i = 0; i < p.ninstr; i++) {
        const FInstr &I = p.ins[i];
        fprintf(stderr, "  %2d %-12s in=%d out=%d logic=%d va=%d vb=%d vout=%d lds=%d vkind=%d store=%d\n", i,
                fkind_name(I.kind), I.fin, I.fout, I.logic, I.va, I.vb, I.vout, I.in_lds, I.vkind, I.store != nullptr);
    }
    for (int j = 0; j < F_MAX_FLAGS; j++)
        if (p.flag_init[j] || p.flag_store[j]) fprintf(stderr, "  flag %d init=%d store=%d\n", j, p.flag_init[j], p.flag_store[j] != nullptr);
}

void KernelBundle::execute(const ExecContext &ctx) const
{
    if (kernels_.empty()) return;
    Engine &eng = *ctx.eng;
    if (eng.tracing) eng.mark("bundle " + (ctx.table ? *ctx.table : std::string("?")));
    if (eng.capture.active && ctx.table && *ctx.table == eng.capture.table) {
        // batched execution: record the fact-table bundle; its program is
        // built once the whole batch is captured (build_captured)
        CapturedBundle cb;
        cb.n = kernels_.front().n;
        cb.segment = ctx.segment;
        cb.kernels = kernels_;
        if (ctx.live) cb.live = *ctx.live;
        cb.table = *ctx.table;
        cb.virgin_begin = ctx.virgin_begin;
        cb.virgin_end = ctx.virgin_end;
        cb.virgin_filled = ctx.virgin_filled;
        eng.capture.bundles.push_back(std::move(cb));
        return;
    }
    if (eng.fusion) {
        FProgram p;
        if (build_program(ctx, p)) {
            validate_program(p, kernels_.front().n, eng.registry);
            dump_program(p, kernels_.front().n, ctx.table ? *ctx.table : std::string("?"));
            const int64_t n = kernels_.front().n;
            // the table's all-true flags, in use after the bundles: a program
            // storing them for every row writes them; one storing a part of
            // them needs the rest filled first (its reads of them stay true)
            bool stores_whole = false, stores_part = false;
            if (ctx.virgin_live && !ctx.virgin_filled)
                for (int j = 0; j < F_MAX_FLAGS; j++) {
                    const uint8_t *st = p.flag_store[j];
                    if (!st || st < ctx.virgin_begin || st >= ctx.virgin_end) continue;
                    if (st == ctx.virgin_begin && n == ctx.virgin_end - ctx.virgin_begin) stores_whole = true;
                    else stores_part = true;
                }
            if (stores_part) {
                eng.registry.check(ctx.virgin_begin, (uint64_t)(ctx.virgin_end - ctx.virgin_begin), "initial flags");
                launch_fill_u8(const_cast<uint8_t *>(ctx.virgin_begin), 1, ctx.virgin_end - ctx.virgin_begin, ctx.stream);
                eng.launches++;
                ctx.virgin_filled = true;
            }
            const bool jit = eng.codegen == SDB_CODEGEN_COMPILE || (eng.codegen == SDB_CODEGEN_AUTO && jit_available());
            // the kernel is chosen, compiled and tuned before the profiling
            // events, which then bracket the program's own launch only
            FusedLaunch L;
            if (jit) L = prepare_fused_jit(p, n, ctx.stream, eng.unroll, eng.codegen == SDB_CODEGEN_COMPILE);
            Engine::ProfRec rec;
            const bool timed = eng.profiles(ctx.table);
            if (timed) eng.prof_begin("fused:" + (ctx.table ? *ctx.table : std::string("?")), n, rec);
            eng.mark("program built");
            if (jit) run_fused_jit(L, p, n, ctx.stream);
            else launch_fused(p, n, ctx.stream);
            if (timed) eng.prof_end(rec);
            eng.launches++;
            if (stores_whole) ctx.virgin_filled = true;
            eng.mark("launched");
            return;
        }
    }
    execute_unfused(ctx);
}

void KernelBundle::build_captured(Engine &eng, CapturedBundle &cb)
{
    KernelBundle b;
    b.kernels_ = cb.kernels;
    ExecContext ctx{&eng, eng.stream, &cb.live, &cb.table};
    ctx.segment = cb.segment;
    ctx.virgin_begin = cb.virgin_begin;
    ctx.virgin_end = cb.virgin_end;
    ctx.virgin_filled = cb.virgin_filled;
    if (!b.build_program(ctx, cb.p)) throw NotBatchable("a fact-table bundle does not fuse");
    try {
        validate_program(cb.p, cb.n, eng.registry);
    } catch (const Error &e) {
        throw NotBatchable(std::string("a fact-table bundle touches a row-sized intermediate: ") + e.what());
    }
    dump_program(cb.p, cb.n, cb.table + " (captured)");
}

void KernelBundle::execute_unfused(const ExecContext &ctx) const
{
    hipStream_t s = ctx.stream;
    Engine &eng = *ctx.eng;
    // a flags buffer that is still logically all-true gets its values now
    // (or one still in use after the bundles: its kernels may write a part)
    bool need_fill = ctx.virgin_live;
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
        const bool timed = eng.profiles(ctx.table);
        if (timed)
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
        if (timed) eng.prof_end(rec);
        eng.launches++;
    }
}

// ------------------------------------------------------------------ fusion
//
// build_program translates the bundle into a fused program in three passes:
//   1. translation: each functor becomes one instruction; every distinct flag
//      buffer becomes a flag register, every column produced inside the bundle
//      a value register (probes use the compact bitmap tables when present);
//   2. optimisation, each step preserving every row's result:
//      - SEARCH triples (x >= lo into a local buffer, x <= hi AND, then the
//        logical merge) fold into one range / two-point predicate;
//      - a NONE predicate whose input is still the all-true initial flags
//        becomes AND (identical on an all-true input);
//      - a run of conjunctive instructions (out = in AND pred(row)) whose
//        intermediate flags nobody else reads is reordered cheapest and most
//        selective first; values a full join produces there are only read
//        under the run's final flag, so rows that reach them are unchanged;
//   3. layout: LDS register file, staged bitmaps, LDS aggregation area, and
//      the write-back of buffers still referenced after the bundle.
// It returns false when the bundle needs something the fused kernel does
// not cover (the bundle then runs unfused).

namespace {

struct Draft {
    FInstr I{};
    float sel = 1.0f;   // fraction of rows a predicate keeps
    float cost = 1.0f;  // relative per-row cost
    bool store = false; // writes a value back to memory
    const void *rank = nullptr; // FK_PROBE_BITS: rank dictionary (bits + values) that may be staged
    uint32_t rank_bytes = 0;
};

// What an FK_PROBE_BITS instruction stages in LDS, if anything: its rank
// dictionary (values in LDS too) or only its presence bitmap.
struct Staging {
    const void *src = nullptr;
    uint32_t bytes = 0;
    bool dict = false;
};

uint32_t bitmap_stage_bytes(const FInstr &I)
{
    const uint64_t len = (uint64_t)((int64_t)I.ht_max - I.ht_min + 1);
    return (uint32_t)(((len + 31) / 32 + 3) / 4 * 16);
}

bool is_predicate(uint8_t k)
{
    return k == FK_SEL_LIT || k == FK_SEL_COL || k == FK_SEL_RANGE || k == FK_SEL_IN2 || k == FK_LOGICAL ||
           k == FK_PROBE_SEMI || k == FK_PROBE_FULL || k == FK_PROBE_BITS;
}

bool writes_flag(uint8_t k) { return is_predicate(k); }

bool is_conjunctive(const FInstr &I)
{
    switch (I.kind) {
    case FK_PROBE_SEMI:
    case FK_PROBE_FULL:
    case FK_PROBE_BITS: return true;
    case FK_SEL_LIT:
    case FK_SEL_COL:
    case FK_SEL_RANGE:
    case FK_SEL_IN2:
    case FK_LOGICAL: return I.logic == SDB_AND;
    default: return false;
    }
}

bool has_vout(const FInstr &I)
{
    switch (I.kind) {
    case FK_FILL:
    case FK_BINOP_CC:
    case FK_BINOP_LC:
    case FK_BINOP_CL:
    case FK_PROBE_FULL: return true;
    case FK_PROBE_BITS: return I.vkind != FV_NONE;
    default: return false;
    }
}

// value registers an instruction reads
void value_reads(const FInstr &I, const FGroup &g, std::vector<int> &out)
{
    out.clear();
    switch (I.kind) {
    case FK_SEL_LIT:
    case FK_SEL_RANGE:
    case FK_SEL_IN2:
    case FK_PROBE_SEMI:
    case FK_PROBE_FULL:
    case FK_PROBE_BITS:
    case FK_BUILD_KEYS:
    case FK_SUM:
    case FK_BINOP_CL:
        if (I.va >= 0) out.push_back(I.va);
        break;
    case FK_SEL_COL:
    case FK_BINOP_CC:
    case FK_BUILD_KEYVALS:
        if (I.va >= 0) out.push_back(I.va);
        if (I.vb >= 0) out.push_back(I.vb);
        break;
    case FK_BINOP_LC:
        if (I.vb >= 0) out.push_back(I.vb);
        break;
    case FK_GROUPBY:
        for (int j = 0; j < g.k; j++)
            if (g.key_reg[j] >= 0) out.push_back(g.key_reg[j]);
        if (g.agg_reg >= 0) out.push_back(g.agg_reg);
        break;
    default: break;
    }
}

// flag registers an instruction reads (fin always, floc for LOGICAL)
int flag_reads(const FInstr &I, int out[2])
{
    int n = 0;
    switch (I.kind) {
    case FK_FILL: return 0; // writes every row, reads no flags
    case FK_LOGICAL: out[n++] = I.fin; out[n++] = I.floc; return n;
    default: out[n++] = I.fin; return n;
    }
}

// SEARCH fold: [x >= lo (NONE) into L][x <= hi AND into L][logic(in, L)]
// or the two-point form [x == a (NONE)][x == b OR][logic(in, L)].
void fold_search(std::vector<Draft> &d)
{
    for (size_t i = 0; i + 2 < d.size(); i++) {
        FInstr &a = d[i].I, &b = d[i + 1].I, &c = d[i + 2].I;
        if (a.kind != FK_SEL_LIT || b.kind != FK_SEL_LIT || c.kind != FK_LOGICAL) continue;
        const int L = a.fout;
        if (a.fin != L || b.fin != L || b.fout != L || c.floc != L || c.fin == L || c.fout == L) continue;
        if (a.va != b.va || a.a != b.a) continue;
        const bool range = a.cmp == SDB_GE && a.logic == SDB_NONE && b.cmp == SDB_LE && b.logic == SDB_AND;
        const bool pts = a.cmp == SDB_EQ && a.logic == SDB_NONE && b.cmp == SDB_EQ && b.logic == SDB_OR;
        if (!range && !pts) continue;
        // the local buffer must not be read after the merge
        bool read_later = false;
        for (size_t k = i + 3; k < d.size() && !read_later; k++) {
            int r[2];
            const int nr = flag_reads(d[k].I, r);
            for (int q = 0; q < nr; q++) read_later |= r[q] == L;
            if (writes_flag(d[k].I.kind) && d[k].I.fout == L) break;
        }
        if (read_later) continue;
        FInstr f{};
        f.kind = range ? FK_SEL_RANGE : FK_SEL_IN2;
        f.logic = c.logic;
        f.fin = c.fin;
        f.fout = c.fout;
        f.va = a.va;
        f.a = a.a;
        f.vb = f.vout = -1;
        f.lit = a.lit;
        f.lit2 = b.lit;
        Draft nd;
        nd.I = f;
        nd.sel = d[i + 2].sel;
        nd.cost = 1.0f;
        d.erase(d.begin() + i, d.begin() + i + 3);
        d.insert(d.begin() + i, nd);
    }
}

// LDS bytes one fused program may give to staged probe tables
// (SDB_STAGE_MAX_BYTES; default 48 KiB): staging trades occupancy (LDS per
// block) for probes that never leave the CU.
uint32_t stage_max_bytes()
{
    static const uint32_t v = [] {
        const char *e = getenv("SDB_STAGE_MAX_BYTES");
        return e && *e ? (uint32_t)atol(e) : 48u * 1024u;
    }();
    return v;
}

// Largest single table staged (SDB_STAGE_TABLE_MAX_BYTES; default 32 KiB).
uint32_t stage_table_max_bytes()
{
    static const uint32_t v = [] {
        const char *e = getenv("SDB_STAGE_TABLE_MAX_BYTES");
        return e && *e ? (uint32_t)atol(e) : 32u * 1024u;
    }();
    return v;
}

// Tables of FK_PROBE_BITS staged into LDS at kernel start, smallest first,
// each at most 32 KiB, together at most 48 KiB: a probe's rank dictionary when
// it has one (its values then come from LDS too), else its presence bitmap;
// a probe whose dictionary does not fit may still stage its bitmap.  Decided
// before the instructions are ordered, since a probe into LDS costs far less
// than one through the caches.  Returns the staging of every instruction.
std::vector<Staging> choose_staged(const std::vector<Draft> &d, uint32_t total_cap)
{
    struct Cand {
        uint32_t bytes;
        const void *src;
        bool dict;
    };
    std::vector<Cand> cand;
    auto add = [&](const void *src, uint32_t bytes, bool dict) {
        for (const Cand &c : cand)
            if (c.src == src) return;
        cand.push_back({bytes, src, dict});
    };
    for (const Draft &D : d) {
        const FInstr &I = D.I;
        if (I.kind != FK_PROBE_BITS) continue;
        if (D.rank && I.vkind == FV_U16) add(D.rank, D.rank_bytes, true);
        const uint64_t len = (uint64_t)((int64_t)I.ht_max - I.ht_min + 1);
        if (len <= (uint64_t)stage_table_max_bytes() * 8u) add(I.ht, bitmap_stage_bytes(I), false);
    }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.bytes < b.bytes; });
    std::vector<const void *> chosen;
    uint32_t staged = 0;
    int nblobs = 0;
    for (const Cand &c : cand) {
        if (c.bytes > stage_table_max_bytes() || nblobs >= F_MAX_BLOBS || staged + c.bytes > total_cap) continue;
        chosen.push_back(c.src);
        staged += c.bytes;
        nblobs++;
    }
    auto has = [&](const void *p) { return p && std::find(chosen.begin(), chosen.end(), p) != chosen.end(); };
    std::vector<Staging> out(d.size());
    for (size_t i = 0; i < d.size(); i++) {
        const Draft &D = d[i];
        if (D.I.kind != FK_PROBE_BITS) continue;
        if (D.rank && D.I.vkind == FV_U16 && has(D.rank)) out[i] = {D.rank, D.rank_bytes, true};
        else if (has(D.I.ht)) out[i] = {D.I.ht, bitmap_stage_bytes(D.I), false};
    }
    return out;
}

// SDB_GROUPBY_HASH=0 sends group-by rows whose dense slot range is too large
// for an LDS array straight to the global arrays instead of an LDS hash.
bool hash_groupby_enabled()
{
    static const bool v = [] {
        const char *e = getenv("SDB_GROUPBY_HASH");
        return !(e && *e == '0');
    }();
    return v;
}

// Columns the kernel prefetches.  A lazily loaded column costs the whole
// wave a dependent memory round trip whenever any of its lanes needs it, which
// for 64 lanes x 4 rows is nearly always, so every column read is prefetched
// unless its 128-byte lines (32 rows) would be touched with probability below
// `min_touch` (estimated from the fraction of rows still needed where it is
// first read).  Past F_MAX_PF columns the most touched win; later reads of a
// prefetched column use its register too.
double prefetch_min_touch()
{
    static const double v = [] {
        const char *e = getenv("SDB_PREFETCH_MIN_TOUCH");
        return e && *e ? atof(e) : 0.3;
    }();
    return v;
}

void assign_prefetch(std::vector<Draft> &d, FGroup &g, FProgram &p)
{
    std::vector<std::pair<const int32_t *, double>> seen; // column, touch probability (first read)
    auto consider = [&](const int32_t *mem, double frac) {
        if (!mem) return;
        for (auto &c : seen)
            if (c.first == mem) return;
        seen.emplace_back(mem, 1.0 - std::pow(1.0 - std::min(1.0, std::max(0.0, frac)), 32.0));
    };
    double live = 1.0;
    for (const Draft &D : d) {
        const FInstr &I = D.I;
        const bool and_input = I.logic == SDB_AND || I.kind == FK_PROBE_SEMI || I.kind == FK_PROBE_FULL ||
                               I.kind == FK_PROBE_BITS || I.kind == FK_BINOP_CC || I.kind == FK_BINOP_LC ||
                               I.kind == FK_BINOP_CL || I.kind == FK_SUM || I.kind == FK_GROUPBY;
        const double frac = and_input ? live : 1.0;
        switch (I.kind) {
        case FK_SEL_COL:
        case FK_BINOP_CC:
        case FK_BUILD_KEYVALS:
            consider(I.a, I.kind == FK_BUILD_KEYVALS ? 1.0 : frac);
            consider(I.b, frac);
            break;
        case FK_BINOP_LC: consider(I.b, frac); break;
        case FK_BUILD_KEYS: consider(I.a, 1.0); break;
        case FK_GROUPBY:
            for (int j = 0; j < g.k; j++) consider(g.key_mem[j], frac);
            consider(g.agg_mem, frac);
            break;
        case FK_LOGICAL:
        case FK_FILL: break;
        default: consider(I.a, frac); break;
        }
        if (is_conjunctive(I)) live *= std::max(0.0f, std::min(1.0f, D.sel));
        else if (is_predicate(I.kind)) live = 1.0;
    }
    std::stable_sort(seen.begin(), seen.end(),
                     [](const std::pair<const int32_t *, double> &a, const std::pair<const int32_t *, double> &b) {
                         return a.second > b.second;
                     });
    std::vector<const int32_t *> pf;
    for (auto &c : seen)
        if (c.second >= prefetch_min_touch() && (int)pf.size() < F_MAX_PF) pf.push_back(c.first);
    p.npf = (int)pf.size();
    for (int j = 0; j < p.npf; j++) p.pf_mem[j] = pf[j];
    auto rewrite = [&](int8_t &reg, const int32_t *&mem) {
        if (reg >= 0 || !mem) return;
        for (int j = 0; j < p.npf; j++)
            if (pf[j] == mem) {
                reg = (int8_t)(F_PF_BASE + j);
                mem = nullptr;
            }
    };
    for (Draft &D : d) {
        FInstr &I = D.I;
        if (I.kind != FK_FILL && I.kind != FK_LOGICAL) {
            if (I.kind != FK_BINOP_LC) rewrite(I.va, I.a);
            if (I.kind == FK_SEL_COL || I.kind == FK_BINOP_CC || I.kind == FK_BINOP_LC || I.kind == FK_BUILD_KEYVALS)
                rewrite(I.vb, I.b);
        }
    }
    for (int j = 0; j < g.k; j++) rewrite(g.key_reg[j], g.key_mem[j]);
    rewrite(g.agg_reg, g.agg_mem);
}

} // namespace

bool KernelBundle::build_program(const ExecContext &ctx, FProgram &p) const
{
    p = FProgram{};
    p.magic = F_MAGIC;
    // a program captured for a batched launch shares the block's LDS with the
    // other pipelines' programs: it gets the capture's budget
    const Engine &eng = *ctx.eng;
    const bool captured = eng.capture.active && ctx.table && *ctx.table == eng.capture.table;
    const uint32_t lds_budget = captured ? eng.capture.lds_budget : F_LDS_BUDGET;
    uint32_t group_area = 0;
    const int64_t n = kernels_.front().n;
    for (const KernelData &k : kernels_)
        if (k.n != n) return false;

    std::vector<const uint8_t *> fptr;     // flag register -> buffer
    std::vector<uint8_t> finit;            // FFlagInit per register
    std::map<const int32_t *, int> vmap;   // column produced in this bundle -> value register
    std::vector<int32_t *> vptr;           // value register -> its column
    bool ok = true;

    auto freg = [&](const uint8_t *ptr, bool write) -> uint8_t {
        for (size_t j = 0; j < fptr.size(); j++)
            if (fptr[j] == ptr) return (uint8_t)j;
        if ((int)fptr.size() >= F_MAX_FLAGS) {
            ok = false;
            return 0;
        }
        fptr.push_back(ptr);
        finit.push_back(write ? FF_UNDEF : (reads_virgin(ctx, ptr) ? FF_TRUE : FF_LOAD));
        return (uint8_t)(fptr.size() - 1);
    };
    auto src = [&](const int32_t *ptr, int8_t &reg, const int32_t *&mem) {
        auto it = vmap.find(ptr);
        if (it != vmap.end()) {
            reg = (int8_t)it->second;
            mem = nullptr;
        } else {
            reg = -1;
            mem = ptr;
        }
    };
    auto dst = [&](int32_t *ptr, FInstr &I, bool &store) {
        auto it = vmap.find(ptr);
        int r;
        if (it != vmap.end()) {
            r = it->second;
        } else {
            r = (int)vptr.size();
            if (r >= F_MAX_VALS) {
                ok = false;
                r = 0;
            }
            vmap[ptr] = r;
            vptr.push_back(ptr);
        }
        I.vout = (int8_t)r;
        store = ctx.live->contains(ptr);
        I.store = store ? ptr : nullptr;
    };

    // ---------------------------------------------------------- translation
    std::vector<Draft> d;
    bool terminal_seen = false;
    for (const KernelData &k0 : kernels_) {
        if (k0.type == KernelType::EmptyKernel) continue;
        if (terminal_seen) return false;
        // a probe's compact tables, with their statistics read
        KernelData k = k0;
        if (k.compact) {
            if (!k.compact->resolved) ctx.eng->resolve_compact();
            if (!k.compact->resolved) fail(SDB_ERR_RUNTIME, "compact table statistics not read");
            k.vals = k.compact->vals;
            k.vkind = k.compact->vkind;
            k.rank = k.compact->rank;
            k.rank_bytes = k.compact->rank_bytes;
            k.sel = k.compact->sel;
        }
        Draft D;
        FInstr &I = D.I;
        I.va = I.vb = I.vout = -1;
        D.sel = k.sel;
        switch (k.type) {
        case KernelType::SelectionKernelLiteral:
        case KernelType::SelectionKernelColumns:
            I.kind = k.type == KernelType::SelectionKernelLiteral ? FK_SEL_LIT : FK_SEL_COL;
            I.cmp = (uint8_t)k.cmp;
            I.logic = (uint8_t)k.logic;
            I.lit = k.literal;
            I.fin = freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            if (I.kind == FK_SEL_COL) src(k.b, I.vb, I.b);
            I.fout = freg(k.out_flags, true);
            D.cost = I.kind == FK_SEL_COL ? 2.0f : 1.0f;
            break;
        case KernelType::LogicalKernel:
            I.kind = FK_LOGICAL;
            I.logic = (uint8_t)k.logic;
            I.fin = freg(k.in_flags, false);
            I.floc = freg(k.local_flags, false);
            I.fout = freg(k.out_flags, true);
            D.cost = 0.1f;
            break;
        case KernelType::FillKernel:
            I.kind = FK_FILL;
            I.lit = k.literal;
            dst(k.result, I, D.store);
            break;
        case KernelType::PerformOperationKernelColumns:
        case KernelType::PerformOperationKernelLiteralFirst:
        case KernelType::PerformOperationKernelLiteralSecond:
            I.kind = k.type == KernelType::PerformOperationKernelColumns        ? FK_BINOP_CC
                     : k.type == KernelType::PerformOperationKernelLiteralFirst ? FK_BINOP_LC
                                                                                 : FK_BINOP_CL;
            I.op = (uint8_t)k.op;
            I.lit = k.literal;
            I.fin = freg(k.in_flags, false);
            if (I.kind != FK_BINOP_LC) src(k.a, I.va, I.a);
            if (I.kind != FK_BINOP_CL) src(k.b, I.vb, I.b);
            dst(k.result, I, D.store);
            break;
        case KernelType::FilterJoinKernel:
        case KernelType::FullJoinKernel: {
            const bool full = k.type == KernelType::FullJoinKernel;
            I.fin = freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            I.ht_min = k.ht_min;
            I.ht_max = k.ht_max;
            if (k.bits) {
                I.kind = FK_PROBE_BITS;
                I.ht = k.bits;
                I.vkind = full ? (int8_t)k.vkind : (int8_t)FV_NONE;
                I.vals = full ? k.vals : nullptr;
                I.vmin = k.vmin;
                if (full && k.rank) {
                    D.rank = k.rank;
                    D.rank_bytes = k.rank_bytes;
                }
                if (full && k.vkind == FV_NONE) return false;
                D.cost = 1.5f;
            } else {
                I.kind = full ? FK_PROBE_FULL : FK_PROBE_SEMI;
                I.ht = k.ht;
                D.cost = full ? 3.0f : 2.0f;
            }
            if (full) {
                if (!aligned(k.ht, 8)) return false;
                dst(k.result, I, D.store);
            }
            I.fout = freg(k.out_flags, true);
            break;
        }
        case KernelType::BuildKeysHTKernel:
            I.kind = FK_BUILD_KEYS;
            I.fin = freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_len = k.ht_len;
            break;
        case KernelType::BuildKeyValsHTKernel:
            I.kind = FK_BUILD_KEYVALS;
            I.fin = freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            src(k.b, I.vb, I.b);
            I.ht = k.ht;
            I.ht_min = k.ht_min;
            I.ht_len = k.ht_len;
            if (!aligned(k.ht, 8)) return false;
            break;
        case KernelType::AggregateOperationKernel:
            I.kind = FK_SUM;
            I.fin = freg(k.in_flags, false);
            src(k.a, I.va, I.a);
            p.sum_acc = k.agg_res;
            if (!aligned(k.agg_res, 8)) return false;
            terminal_seen = true;
            break;
        case KernelType::GroupByAggregateKernel: {
            const GroupBySpec &g = *k.gb;
            const int nk = (int)g.keys.size();
            if (nk < 1 || nk > F_MAX_KEYS || g.prod == 0 || !aligned(g.sums, 8)) return false;
            I.kind = FK_GROUPBY;
            I.fin = freg(k.in_flags, false);
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
            // in-range keys map one-to-one onto slots below prod <= 2^31 (so
            // the reference's int32 slot is non-negative and needs no
            // modulo): LDS copies (dense for small ranges, hashed otherwise)
            // with an exact flush
            fits = fits && exact == g.prod && g.prod <= 0x80000000ull;
            G.lds = !fits ? FG_GLOBAL : (g.prod <= F_DENSE_MAX ? FG_DENSE_LDS : FG_HASH_LDS);
            if (G.lds == FG_HASH_LDS && !hash_groupby_enabled()) G.lds = FG_GLOBAL;
            group_area = G.lds == FG_DENSE_LDS ? (uint32_t)(g.prod * 8 + (g.prod + 31) / 32 * 4 + 16)
                         : G.lds == FG_HASH_LDS ? F_HASH_CAP * 12u + 16u
                                                : 0u;
            if (group_area > lds_budget) {
                G.lds = FG_GLOBAL;
                group_area = 0;
            }
            terminal_seen = true;
            break;
        }
        default: return false;
        }
        if (!ok) return false;
        d.push_back(D);
    }
    if (d.empty() || (int)d.size() > F_MAX_INSTR + 8) return false;

    // --------------------------------------------------------- optimisation
    fold_search(d);
    // NONE on the untouched all-true initial flags is AND
    {
        std::vector<bool> written(fptr.size(), false);
        for (Draft &D : d) {
            FInstr &I = D.I;
            if (is_predicate(I.kind) && I.kind != FK_PROBE_SEMI && I.kind != FK_PROBE_FULL && I.kind != FK_PROBE_BITS &&
                I.logic == SDB_NONE && finit[I.fin] == FF_TRUE && !written[I.fin]) {
                I.logic = SDB_AND;
            }
            if (writes_flag(I.kind)) written[I.fout] = true;
        }
    }
    // LDS left for staged probe tables.  A program captured for a batched
    // launch stages none: the batch decides its probes on fact key columns
    // through shared masks, and its LDS share goes to the group-by partials.
    // (the interpreter's value-register file, 4 KiB per register, shares
    // the budget)
    const uint32_t reserved = group_area + (uint32_t)vptr.size() * 256u * 16u;
    const uint32_t stage_cap =
        captured ? 0u : std::min<uint32_t>(stage_max_bytes(), lds_budget - std::min(lds_budget, reserved));
    // probe costs follow where the tables will live
    {
        const std::vector<Staging> st = choose_staged(d, stage_cap);
        for (size_t i = 0; i < d.size(); i++) {
            Draft &D = d[i];
            if (D.I.kind != FK_PROBE_BITS) continue;
            const bool lds_bits = st[i].src != nullptr;
            const bool lds_vals = D.I.vkind == FV_NONE || st[i].dict;
            D.cost = lds_bits ? (lds_vals ? 0.6f : 1.5f) : 2.0f;
        }
    }
    // reorder conjunctive runs
    {
        const size_t m = d.size();
        // producer of each flag read: index of the last earlier writer of that register, -1 initial
        std::vector<std::array<int, 2>> prod_of(m, {-2, -2});
        std::vector<std::vector<int>> readers(m); // readers of the version instruction i writes
        std::vector<int> last(fptr.size(), -1);
        for (size_t i = 0; i < m; i++) {
            int r[2];
            const int nr = flag_reads(d[i].I, r);
            for (int q = 0; q < nr; q++) {
                const int w = last[r[q]];
                prod_of[i][q] = w;
                if (w >= 0) readers[w].push_back((int)i);
            }
            if (writes_flag(d[i].I.kind)) last[d[i].I.fout] = (int)i;
        }
        auto live_at_end = [&](int i) {
            const int reg = d[i].I.fout;
            return last[reg] == i && ctx.live->contains(fptr[reg]);
        };
        size_t s0 = 0;
        while (s0 < m) {
            if (!is_conjunctive(d[s0].I)) {
                s0++;
                continue;
            }
            size_t e = s0;
            while (e + 1 < m && is_conjunctive(d[e + 1].I) && prod_of[e + 1][0] == (int)e &&
                   readers[e].size() == 1 && !live_at_end((int)e))
                e++;
            const size_t len = e - s0 + 1;
            bool can = len >= 2;
            for (size_t i = s0; i <= e && can; i++)
                if (d[i].I.kind == FK_LOGICAL && prod_of[i][1] >= (int)s0) can = false;
            for (size_t i = s0; i <= e && can; i++) can = !d[i].store;
            // values produced in the run: readers after the run must use its final flag
            for (size_t i = s0; i <= e && can; i++) {
                if (!has_vout(d[i].I)) continue;
                std::vector<int> vr;
                for (size_t k = e + 1; k < m && can; k++) {
                    value_reads(d[k].I, p.gb, vr);
                    bool reads = false;
                    for (int r : vr) reads |= r == d[i].I.vout;
                    if (reads && prod_of[k][0] != (int)e) can = false;
                    if (has_vout(d[k].I) && d[k].I.vout == d[i].I.vout) break;
                }
            }
            if (can) {
                // greedy order by cost / (1 - sel), respecting value dependencies
                std::vector<Draft> run(d.begin() + s0, d.begin() + e + 1);
                std::vector<bool> placed(len, false);
                std::vector<Draft> ordered;
                const uint8_t in0 = run[0].I.fin, out_final = run[len - 1].I.fout;
                for (size_t step = 0; step < len; step++) {
                    int best = -1;
                    float best_rank = 0;
                    for (size_t c = 0; c < len; c++) {
                        if (placed[c]) continue;
                        // ready when every value it reads that the run produces is placed
                        std::vector<int> vr;
                        value_reads(run[c].I, p.gb, vr);
                        bool ready = true;
                        for (size_t q = 0; q < len && ready; q++) {
                            if (placed[q] || q == c || !has_vout(run[q].I)) continue;
                            for (int r : vr) ready &= r != run[q].I.vout;
                        }
                        if (!ready) continue;
                        const float keep = std::min(run[c].sel, 0.9999f);
                        const float rank = run[c].cost / (1.0f - keep);
                        if (best < 0 || rank < best_rank) {
                            best = (int)c;
                            best_rank = rank;
                        }
                    }
                    if (best < 0) {
                        can = false;
                        break;
                    }
                    placed[best] = true;
                    ordered.push_back(run[best]);
                }
                if (can) {
                    for (size_t q = 0; q < len; q++) {
                        ordered[q].I.fin = q == 0 ? in0 : out_final;
                        ordered[q].I.fout = out_final;
                        d[s0 + q] = ordered[q];
                    }
                }
            }
            s0 = e + 1;
        }
    }

    assign_prefetch(d, p.gb, p);

    // --------------------------------------------------------------- layout
    if ((int)d.size() > F_MAX_INSTR) return false;
    bool vec = true;
    for (int j = 0; j < p.npf; j++) vec &= aligned(p.pf_mem[j], 16);
    for (size_t i = 0; i < d.size(); i++) {
        const FInstr &I = d[i].I;
        for (const int32_t *mem : {I.a, I.b})
            if (mem) vec &= aligned(mem, 16);
        if (I.store) vec &= aligned(I.store, 16);
        p.ins[i] = I;
        p.ins[i].sel = d[i].sel;
    }
    p.ninstr = (int)d.size();
    for (int j = 0; j < p.gb.k; j++)
        if (p.gb.key_mem[j]) vec &= aligned(p.gb.key_mem[j], 16);
    if (p.gb.agg_mem) vec &= aligned(p.gb.agg_mem, 16);

    uint32_t off = 0;
    auto take = [&](uint32_t bytes) {
        const uint32_t at = off;
        off += (bytes + 15u) & ~15u;
        return at;
    };
    p.nvals = (int)vptr.size();
    // staged tables (chosen again on the final order: the choice depends only
    // on the set of probes, not on their order)
    {
        const std::vector<Staging> st = choose_staged(d, stage_cap);
        for (int i = 0; i < p.ninstr; i++) {
            FInstr &I = p.ins[i];
            if (!st[i].src) continue;
            int blob = -1;
            for (int b = 0; b < p.nblobs; b++)
                if (p.blob_src[b] == st[i].src) blob = b;
            if (blob < 0) {
                if (p.nblobs >= F_MAX_BLOBS) return false;
                blob = p.nblobs++;
                p.blob_src[blob] = st[i].src;
                p.blob_bytes[blob] = st[i].bytes;
                p.blob_off[blob] = take(st[i].bytes);
            }
            I.in_lds = 1;
            I.lds_off = p.blob_off[blob];
            I.lds_bytes = st[i].bytes;
            if (st[i].dict) {
                I.vkind = FV_RANK16;
                I.vals = nullptr;
                I.ht_len = (int32_t)((((int64_t)I.ht_max - I.ht_min + 1) + 31) / 32);
            }
        }
    }
    if (p.gb.lds == FG_DENSE_LDS) {
        p.gb.lds_off = take((uint32_t)(p.gb.prod * 8 + (p.gb.prod + 31) / 32 * 4));
    } else if (p.gb.lds == FG_HASH_LDS) {
        p.gb.hash_cap = F_HASH_CAP;
        p.gb.lds_off = take(F_HASH_CAP * 12u);
    }
    // the interpreter's value-register file comes last: a kernel compiled
    // for the program keeps its registers in VGPRs and needs only the LDS
    // below vreg_off
    p.vreg_off = take((uint32_t)p.nvals * 256u * 16u);
    p.lds_bytes = off;
    if (p.lds_bytes > F_LDS_BUDGET) return false;
    if (captured && p.vreg_off > lds_budget) return false;

    // flags: initial values and write-back of changed buffers still referenced
    std::vector<bool> written(fptr.size(), false);
    std::vector<int> first(fptr.size(), 0); // 1: first access reads, 2: first access writes
    for (int i = 0; i < p.ninstr; i++) {
        int r[2];
        const int nr = flag_reads(p.ins[i], r);
        for (int q = 0; q < nr; q++)
            if (!first[r[q]]) first[r[q]] = 1;
        if (writes_flag(p.ins[i].kind)) {
            if (!first[p.ins[i].fout]) first[p.ins[i].fout] = 2;
            written[p.ins[i].fout] = true;
        }
    }
    for (size_t j = 0; j < fptr.size(); j++) {
        // a register whose initial value nobody reads is not loaded (an
        // all-true buffer still referenced afterwards is written back true)
        const bool keep_true = finit[j] == FF_TRUE && !written[j] && ctx.live->contains(fptr[j]);
        if (first[j] != 1 && !keep_true) finit[j] = FF_UNDEF;
        p.flag_init[j] = finit[j];
        if (finit[j] == FF_LOAD) {
            p.flag_load[j] = fptr[j];
            vec &= aligned(fptr[j], 4);
        }
        const bool changed = written[j] || finit[j] == FF_TRUE;
        if (changed && ctx.live->contains(fptr[j])) {
            p.flag_store[j] = const_cast<uint8_t *>(fptr[j]);
            vec &= aligned(fptr[j], 4);
        }
    }
    p.vec = vec ? 1 : 0;
    return true;
}

} // namespace sdb
