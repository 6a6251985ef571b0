// TransientTable on one HIP device: models/transient_table.hpp:20-1928 with
// models/models.hpp's operator factories folded in.  Flags are double
// buffered (flags_ / flags_secondary_), operators are queued per segment
// into pending bundles that merge into the previous phase
// (add_pending_kernels, 347-381), and bundles are flushed where the reference
// flushes them: at aggregates, on the build side of joins, and at the end of
// the plan.
#include <cstring>
#include <iostream>

#include "engine.hpp"

namespace sdb {

namespace {

int get_comp_op(const std::string &op) { return sdb_comp_op_from_string(op.c_str()); }
int get_logical_op(const std::string &op) { return sdb_logical_op_from_string(op.c_str()); }

int get_binop(const std::string &op)
{
    const int b = sdb_binop_from_string(op.c_str());
    if (b < 0) throw std::invalid_argument("Unknown operation: " + op); // aggregation.hpp:47
    return b;
}

// is_filter_logical, kernels/selection.hpp:53-59
bool is_filter_logical(const std::string &op)
{
    for (char c : op)
        if (!isalpha((unsigned char)c)) return false;
    return true;
}

// (int) of an int64 literal
int32_t to_int32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }

float clamp01(double x) { return (float)(x < 0 ? 0 : (x > 1 ? 1 : x)); }

// SDB_JOIN_STATS=1: joins read their compact tables' statistics (one stream
// sync per plan) instead of taking the build side's estimated selectivity
// and bounds.
bool join_estimates()
{
    static const bool on = [] {
        const char *e = getenv("SDB_JOIN_STATS");
        return !(e && *e == '1');
    }();
    return on;
}

// Fraction of a uniformly distributed column in [mn, mx] a comparison keeps
// (used only to order fused predicates, never for results).
float estimate_cmp(int cmp, int32_t lit, std::pair<int32_t, int32_t> mm)
{
    const double lo = mm.first, hi = mm.second, range = hi - lo + 1, v = lit;
    switch (cmp) {
    case SDB_EQ: return (v >= lo && v <= hi) ? clamp01(1.0 / range) : 0.0f;
    case SDB_NE: return (v >= lo && v <= hi) ? clamp01(1.0 - 1.0 / range) : 1.0f;
    case SDB_LT: return clamp01((v - lo) / range);
    case SDB_LE: return clamp01((v - lo + 1) / range);
    case SDB_GT: return clamp01((hi - v) / range);
    default: return clamp01((hi - v + 1) / range);
    }
}

} // namespace

std::pair<int32_t, int32_t> Column::get_min_max() const
{
    if (is_aggregate) throw std::runtime_error("can't get min/max of aggregate result column");
    if (seg_min.empty()) throw std::runtime_error("min/max of an empty column");
    int32_t mn = seg_min[0], mx = seg_max[0];
    for (size_t i = 1; i < seg_min.size(); i++) {
        mn = std::min(mn, seg_min[i]);
        mx = std::max(mx, seg_max[i]);
    }
    return {mn, mx};
}

TransientTable::TransientTable(Engine &eng, Table *base) : eng_(eng), nrows_(base->nrows), name_(base->name)
{
    // flags start true (transient_table.hpp:84-106); the fill is deferred:
    // the buffer stays logically all-true until an operator reads it
    flags_ = eng_.row_buffer<uint8_t>(name_, std::max<uint64_t>(nrows_, 1));
    virgin_ = flags_;
    current_columns_.reserve(base->columns.size() + 64);
    for (Column &c : base->columns) current_columns_.push_back(&c);
}

void TransientTable::set_group_by_column(uint64_t col)
{
    if (col >= current_columns_.size()) throw std::invalid_argument("group-by column out of range");
    group_by_column_ = current_columns_[col];
    group_by_column_index_ = col;
}

Column *TransientTable::col(int64_t idx, const char *what) const
{
    if (idx < 0 || idx >= (int64_t)current_columns_.size())
        throw std::invalid_argument(std::string(what) + ": column index " + std::to_string(idx) + " out of range");
    Column *c = current_columns_[idx];
    if (!c) throw std::runtime_error(std::string(what) + ": column " + std::to_string(idx) + " is not available");
    if (!c->loaded) throw std::runtime_error(std::string(what) + ": column " + std::to_string(idx) + " is not loaded");
    return c;
}

Column &TransientTable::new_column(uint64_t nrows, bool aggregate)
{
    Column &c = materialized_columns_.emplace_back();
    c.nrows = nrows;
    c.segment_size = eng_.segment_size;
    c.loaded = true;
    c.is_aggregate = aggregate;
    c.seg_min.assign(c.nseg(), 0);
    c.seg_max.assign(c.nseg(), 0);
    return c;
}

void TransientTable::ensure_secondary_flags()
{
    if (!flags_secondary_) flags_secondary_ = eng_.row_buffer<uint8_t>(name_, std::max<uint64_t>(nrows_, 1));
}

void TransientTable::swap_flag_buffers()
{
    if (flags_secondary_) std::swap(flags_, flags_secondary_);
}

void TransientTable::add_pending_kernels(std::vector<KernelBundle> ops)
{
    if (pending_kernels_.empty() || pending_kernels_.back().size() != ops.size()) {
        pending_kernels_.push_back(std::move(ops));
        return;
    }
    auto &last = pending_kernels_.back();
    for (size_t i = 0; i < ops.size(); i++)
        for (const KernelData &k : ops[i].get_kernels()) last[i].add_kernel(k);
}

void TransientTable::check_flags_memory() const
{
    try {
        eng_.registry.check(flags_, std::max<uint64_t>(nrows_, 1), "initial flags");
    } catch (const Error &) {
        // a row-sized placeholder of a captured fact table
        if (eng_.capture.active && name_ == eng_.capture.table)
            throw NotBatchable("the fact table's flags are needed in memory");
        throw;
    }
}

const uint8_t *TransientTable::flags_device()
{
    if (virgin_ == flags_) {
        check_flags_memory();
        launch_fill_u8(flags_, 1, (int64_t)nrows_, eng_.stream);
        eng_.launches++;
        virgin_ = nullptr;
    }
    return flags_;
}

void TransientTable::execute_pending_kernels(bool dead_after)
{
    if (pending_kernels_.empty()) return;
    LiveSet live;
    bool virgin_live = false;
    if (!dead_after) {
        // flags and columns the table still exposes must hold their values;
        // flags still all true are written by a program that stores them
        // whole, else filled (KernelBundle::execute, and below)
        if (virgin_ == flags_) {
            check_flags_memory();
            virgin_live = true;
        }
        live.add(flags_, nrows_);
        for (Column *c : current_columns_)
            if (c && c->loaded && c->data) live.add(c->data, c->nrows * (c->is_aggregate ? 8 : 4));
        if (group_by_column_ && group_by_column_->data)
            live.add(group_by_column_->data, group_by_column_->nrows * 4);
    }
    ExecContext ctx{&eng_, eng_.stream, &live, &name_};
    if (virgin_) {
        ctx.virgin_begin = virgin_;
        ctx.virgin_end = virgin_ + nrows_;
    }
    ctx.virgin_live = virgin_live;
    const uint64_t nseg = segment_num();
    for (const auto &phase : pending_kernels_)
        if (phase.size() != nseg) throw std::runtime_error("Pending kernels segment number mismatch.");
    for (uint64_t s = 0; s < nseg; s++) {
        ctx.segment = s;
        for (const auto &phase : pending_kernels_) phase[s].execute(ctx);
    }
    if (virgin_live && !ctx.virgin_filled) { // no bundle wrote them
        launch_fill_u8(flags_, 1, (int64_t)nrows_, eng_.stream);
        eng_.launches++;
    }
    pending_kernels_.clear();
    virgin_ = nullptr;
}

// -------------------------------------------------------------------- filter

// The kept fraction after a filter that combines its own estimate with the
// table's flags by `logic` (predicates taken as independent).
void TransientTable::note_filter(int logic, float sel)
{
    if (logic == SDB_AND) est_kept_ *= sel;
    else if (logic == SDB_OR) est_kept_ = est_kept_ + sel - est_kept_ * sel;
    else est_kept_ = sel;
}

void TransientTable::apply_filter(const ExprType &expr, const std::string &parent_op)
{
    if (expr.exprType != ExprOption::EXPR) {
        std::cerr << "Filter condition: Unsupported parsing ExprType" << std::endl;
        return;
    }
    if (expr.op == "SEARCH") {
        apply_search(expr, parent_op);
        return;
    }
    if (is_filter_logical(expr.op)) {
        // left fold: the parent op goes to the first operand, this op to the rest
        bool used = false;
        for (const ExprType &operand : expr.operands) {
            apply_filter(operand, used ? expr.op : parent_op);
            used = true;
        }
        return;
    }
    if (expr.operands.size() != 2) {
        std::cerr << "Filter condition: Unsupported number of operands for EXPR" << std::endl;
        return;
    }
    Column *cols[2] = {nullptr, nullptr};
    bool literal = false;
    int32_t literal_value = 0;
    for (int i = 0; i < 2; i++) {
        const ExprType &o = expr.operands[i];
        if (o.exprType == ExprOption::COLUMN) {
            cols[i] = col(o.input, "filter");
        } else if (o.exprType == ExprOption::LITERAL) {
            literal = true;
            literal_value = to_int32(o.literal.value);
        } else {
            std::cerr << "Filter condition: Unsupported parsing ExprType for comparison operand" << std::endl;
            return;
        }
    }
    if (!cols[0]) throw std::invalid_argument("filter: the first comparison operand must be a column");
    ensure_secondary_flags();
    const int cmp = get_comp_op(expr.op), logic = get_logical_op(parent_op);
    const float sel = literal ? estimate_cmp(cmp, literal_value, cols[0]->get_min_max()) : 0.5f;
    note_filter(logic, sel);
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        const uint64_t off = s * eng_.segment_size;
        KernelData k;
        k.type = literal ? KernelType::SelectionKernelLiteral : KernelType::SelectionKernelColumns;
        k.n = (int64_t)seg_rows(s);
        k.cmp = cmp;
        k.logic = logic;
        k.in_flags = flags_ + off;
        k.out_flags = flags_secondary_ + off;
        k.a = cols[0]->i32(s);
        if (literal) k.literal = literal_value;
        else k.b = cols[1]->i32(s);
        k.sel = sel;
        ops[s].add_kernel(k);
    }
    add_pending_kernels(std::move(ops));
    swap_flag_buffers();
}

// SEARCH: Segment::search_operator, models/models.hpp:604-718.  One range
// reads rangeSet[0][1..2]; otherwise two points rangeSet[0][1], [1][1].
void TransientTable::apply_search(const ExprType &expr, const std::string &parent_op)
{
    if (expr.operands.size() < 2) throw std::invalid_argument("SEARCH needs two operands");
    Column *c = col(expr.operands[0].input, "search");
    const auto &rs = expr.operands[1].literal.rangeSet;
    int32_t v0, v1;
    const bool range = rs.size() == 1;
    if (range) {
        if (rs[0].size() < 3) throw std::invalid_argument("SEARCH range needs [kind, lower, upper]");
        v0 = std::stoi(rs[0][1]);
        v1 = std::stoi(rs[0][2]);
    } else {
        if (rs.size() < 2 || rs[0].size() < 2 || rs[1].size() < 2)
            throw std::invalid_argument("SEARCH needs one range or two points");
        v0 = std::stoi(rs[0][1]);
        v1 = std::stoi(rs[1][1]);
    }
    ensure_secondary_flags();
    const int logic = get_logical_op(parent_op);
    float sel;
    {
        const auto mm = c->get_min_max();
        const double width = (double)mm.second - mm.first + 1;
        sel = range ? clamp01(((double)std::min(v1, mm.second) - std::max(v0, mm.first) + 1) / width)
                    : clamp01(2.0 / width);
    }
    note_filter(logic, sel);
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        const uint64_t off = s * eng_.segment_size;
        const int64_t n = (int64_t)seg_rows(s);
        uint8_t *local = eng_.row_buffer<uint8_t>(name_, n);
        KernelData k1, k2, k3;
        k1.type = k2.type = KernelType::SelectionKernelLiteral;
        k1.n = k2.n = k3.n = n;
        k1.in_flags = k2.in_flags = local;
        k1.out_flags = k2.out_flags = local;
        k1.a = k2.a = c->i32(s);
        if (range) {
            k1.cmp = SDB_GE, k1.logic = SDB_NONE, k1.literal = v0;
            k2.cmp = SDB_LE, k2.logic = SDB_AND, k2.literal = v1;
        } else {
            k1.cmp = SDB_EQ, k1.logic = SDB_NONE, k1.literal = v0;
            k2.cmp = SDB_EQ, k2.logic = SDB_OR, k2.literal = v1;
        }
        k3.type = KernelType::LogicalKernel;
        k3.logic = logic;
        k3.sel = sel;
        k3.in_flags = flags_ + off;
        k3.local_flags = local;
        k3.out_flags = flags_secondary_ + off;
        ops[s].add_kernel(k1);
        ops[s].add_kernel(k2);
        ops[s].add_kernel(k3);
    }
    add_pending_kernels(std::move(ops));
    swap_flag_buffers();
}

// ------------------------------------------------------------------- project

void TransientTable::apply_project(const std::vector<ExprType> &exprs)
{
    std::vector<Column *> new_columns;
    new_columns.reserve(exprs.size() + 50);
    for (size_t i = 0; i < exprs.size(); i++) {
        const ExprType &e = exprs[i];
        switch (e.exprType) {
        case ExprOption::COLUMN: {
            if (e.input < 0 || e.input >= (int64_t)current_columns_.size())
                throw std::invalid_argument("project: column index out of range");
            new_columns.push_back(current_columns_[e.input]);
            if ((uint64_t)e.input == group_by_column_index_) group_by_column_index_ = i;
            break;
        }
        case ExprOption::LITERAL: {
            const int32_t v = to_int32(e.literal.value);
            Column &nc = new_column(nrows_);
            nc.data = eng_.row_buffer<int32_t>(name_, std::max<uint64_t>(nrows_, 1));
            std::vector<KernelBundle> ops(segment_num());
            for (uint64_t s = 0; s < segment_num(); s++) {
                KernelData k;
                k.type = KernelType::FillKernel;
                k.n = (int64_t)seg_rows(s);
                k.result = nc.i32(s);
                k.literal = v;
                nc.seg_min[s] = nc.seg_max[s] = v; // models.hpp:336-337
                ops[s].add_kernel(k);
            }
            add_pending_kernels(std::move(ops));
            new_columns.push_back(&nc);
            break;
        }
        case ExprOption::EXPR: {
            if (e.operands.size() != 2) {
                std::cerr << "Project operation: Unsupported number of operands for EXPR" << std::endl;
                return;
            }
            const ExprOption k0 = e.operands[0].exprType, k1 = e.operands[1].exprType;
            KernelType kt;
            if (k0 == ExprOption::COLUMN && k1 == ExprOption::COLUMN) kt = KernelType::PerformOperationKernelColumns;
            else if (k0 == ExprOption::LITERAL && k1 == ExprOption::COLUMN) kt = KernelType::PerformOperationKernelLiteralFirst;
            else if (k0 == ExprOption::COLUMN && k1 == ExprOption::LITERAL) kt = KernelType::PerformOperationKernelLiteralSecond;
            else {
                std::cerr << "Project operation: Unsupported parsing ExprType for EXPR" << std::endl;
                return;
            }
            const int op = get_binop(e.op);
            Column *ca = k0 == ExprOption::COLUMN ? col(e.operands[0].input, "project") : nullptr;
            Column *cb = k1 == ExprOption::COLUMN ? col(e.operands[1].input, "project") : nullptr;
            Column &nc = new_column(nrows_);
            nc.data = eng_.row_buffer<int32_t>(name_, std::max<uint64_t>(nrows_, 1));
            std::vector<KernelBundle> ops(segment_num());
            for (uint64_t s = 0; s < segment_num(); s++) {
                KernelData k;
                k.type = kt;
                k.n = (int64_t)seg_rows(s);
                k.op = op;
                k.result = nc.i32(s);
                k.in_flags = flags_ + s * eng_.segment_size;
                if (ca) k.a = ca->i32(s);
                if (cb) k.b = cb->i32(s);
                if (kt == KernelType::PerformOperationKernelLiteralFirst) k.literal = to_int32(e.operands[0].literal.value);
                if (kt == KernelType::PerformOperationKernelLiteralSecond) k.literal = to_int32(e.operands[1].literal.value);
                // segment min/max: models/models.hpp:774-775, 803-804, 832-833
                if (ca && cb) {
                    nc.seg_min[s] = std::min(ca->seg_min[s], cb->seg_min[s]);
                    nc.seg_max[s] = std::max(ca->seg_max[s], cb->seg_max[s]);
                } else {
                    Column *c = ca ? ca : cb;
                    nc.seg_min[s] = c->seg_min[s];
                    nc.seg_max[s] = c->seg_max[s];
                }
                ops[s].add_kernel(k);
            }
            add_pending_kernels(std::move(ops));
            new_columns.push_back(&nc);
            break;
        }
        }
    }
    current_columns_ = std::move(new_columns);
}

// ----------------------------------------------------------------- aggregate

void TransientTable::apply_aggregate(const AggType &agg, const std::vector<int64_t> &group)
{
    if (agg.operands.empty()) throw std::invalid_argument("aggregate without operand");
    Column *agg_col = col(agg.operands[0], "aggregate");
    if (group.empty()) {
        // scalar: transient_table.hpp:1234-1380
        uint64_t *result = eng_.mm->alloc_zero<uint64_t>(1);
        std::vector<KernelBundle> ops(segment_num());
        for (uint64_t s = 0; s < segment_num(); s++) {
            KernelData k;
            k.type = KernelType::AggregateOperationKernel;
            k.n = (int64_t)seg_rows(s);
            k.a = agg_col->i32(s);
            k.in_flags = flags_ + s * eng_.segment_size;
            k.agg_res = result;
            ops[s].add_kernel(k);
        }
        add_pending_kernels(std::move(ops));
        execute_pending_kernels(true);
        nrows_ = 1;
        flags_ = eng_.mm->alloc<uint8_t>(1);
        virgin_ = flags_;
        flags_secondary_ = nullptr;
        Column &r = new_column(1, true);
        r.data = result;
        current_columns_.clear();
        current_columns_.push_back(&r);
        return;
    }
    // group-by: transient_table.hpp:1381-1646
    std::vector<Column *> keys;
    std::vector<int32_t> mins, maxs;
    uint64_t prod = 1;
    for (int64_t g : group) {
        Column *c = col(g, "group-by");
        auto mm = c->get_min_max();
        keys.push_back(c);
        mins.push_back(mm.first);
        maxs.push_back(mm.second);
        // prod_ranges *= max - min + 1: an int expression widened to uint64
        prod *= (uint64_t)(int64_t)(int32_t)((uint32_t)mm.second - (uint32_t)mm.first + 1u);
    }
    if (prod == 0 || prod > (1ull << 34)) throw std::bad_alloc();
    uint64_t *sums = eng_.mm->alloc_zero<uint64_t>(prod);
    uint32_t *seen = eng_.mm->alloc_zero<uint32_t>(prod);
    std::vector<int32_t *> out_keys;
    for (size_t j = 0; j < keys.size(); j++) out_keys.push_back(eng_.mm->alloc<int32_t>(prod));
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        auto spec = std::make_shared<GroupBySpec>();
        for (Column *c : keys) spec->keys.push_back(c->i32(s));
        spec->mins = mins;
        spec->maxs = maxs;
        spec->out_keys = out_keys;
        spec->sums = sums;
        spec->seen = seen;
        spec->prod = prod;
        KernelData k;
        k.type = KernelType::GroupByAggregateKernel;
        k.n = (int64_t)seg_rows(s);
        k.a = agg_col->i32(s);
        k.in_flags = flags_ + s * eng_.segment_size;
        k.gb = spec;
        ops[s].add_kernel(k);
    }
    add_pending_kernels(std::move(ops));
    execute_pending_kernels(true);
    nrows_ = prod;
    flags_ = eng_.mm->alloc<uint8_t>(prod);
    if (eng_.capture.active && name_ == eng_.capture.table) {
        // the fact pass has not run yet: convert after the batched launch
        uint8_t *f = flags_;
        Engine *eng = &eng_;
        eng_.capture.deferred.push_back([=] {
            launch_seen_to_flags(seen, f, prod, eng->stream);
            eng->launches++;
        });
    } else {
        launch_seen_to_flags(seen, flags_, prod, eng_.stream);
        eng_.launches++;
    }
    virgin_ = nullptr;
    flags_secondary_ = nullptr;
    current_columns_.clear();
    for (int32_t *k : out_keys) {
        Column &c = new_column(prod); // key columns carry min = max = 0 (models.hpp:182-183)
        c.data = k;
        current_columns_.push_back(&c);
    }
    Column &sc = new_column(prod, true);
    sc.data = sums;
    current_columns_.push_back(&sc);
}

// --------------------------------------------------------------------- joins

std::tuple<uint8_t *, int32_t, int32_t> TransientTable::build_keys_hash_table(int column)
{
    // Column::build_keys_hash_table, models/models.hpp:1413-1452
    Column *c = col(column, "join build");
    auto mm = c->get_min_max();
    const int32_t len = (int32_t)((uint32_t)mm.second - (uint32_t)mm.first + 1u);
    if (len <= 0) throw std::runtime_error("join build: key range does not fit an int");
    uint8_t *ht = eng_.mm->alloc_zero<uint8_t>((uint64_t)len);
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        KernelData k;
        k.type = KernelType::BuildKeysHTKernel;
        k.n = (int64_t)seg_rows(s);
        k.a = c->i32(s);
        k.in_flags = flags_ + s * eng_.segment_size;
        k.ht = ht;
        k.ht_len = len;
        k.ht_min = mm.first;
        ops[s].add_kernel(k);
    }
    add_pending_kernels(std::move(ops));
    return {ht, mm.first, mm.second};
}

std::tuple<int32_t *, int32_t, int32_t> TransientTable::build_key_vals_hash_table(int column)
{
    // Column::build_key_vals_hash_table, models/models.hpp:1509-1548
    Column *c = col(column, "join build");
    if (!group_by_column_) throw std::runtime_error("full join: the build table has no group-by column");
    Column *vals = group_by_column_;
    auto mm = c->get_min_max();
    const int32_t len = (int32_t)((uint32_t)mm.second - (uint32_t)mm.first + 1u);
    if (len <= 0) throw std::runtime_error("join build: key range does not fit an int");
    int32_t *ht = eng_.mm->alloc_zero<int32_t>(2 * (uint64_t)len);
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        KernelData k;
        k.type = KernelType::BuildKeyValsHTKernel;
        k.n = (int64_t)seg_rows(s);
        k.a = c->i32(s);
        k.b = vals->i32(s);
        k.in_flags = flags_ + s * eng_.segment_size;
        k.ht = ht;
        k.ht_len = len;
        k.ht_min = mm.first;
        ops[s].add_kernel(k);
    }
    add_pending_kernels(std::move(ops));
    return {ht, mm.first, mm.second};
}

// Compact probe tables for fused bundles, derived from the reference-layout
// table once the build side ran: a presence bitmap (and 16-bit value offsets
// for a full join when the values fit, with their rank dictionary when small
// enough for LDS), plus the fraction of set slots, which orders predicates.
// Results never depend on them: a bit is set exactly when the reference
// table reports a match.  Only the launches happen here; the statistics are
// read on first use (Engine::resolve_compact).
// The value layout of a pair table, from its number of set slots (or a bound
// on it) and its overflow flag: the 16-bit values when they fit, with their
// rank dictionary staged for `set` values when that fits its buffer.
static void settle_values(CompactState &c, uint32_t set, uint32_t overflow)
{
    if (c.narrow && overflow == 0) {
        c.vals = c.narrow;
        c.vkind = FV_U16;
        const uint32_t dict_bytes = (rank_values_off((uint32_t)c.words) + 2u * set + 15u) / 16u * 16u;
        if (c.dict && set <= 0xFFFFu && dict_bytes <= c.dict_cap) {
            c.rank = c.dict;
            c.rank_bytes = dict_bytes;
        }
    } else {
        c.vals = c.ht;
        c.vkind = FV_PAIR;
    }
}

// With `est_sel` >= 0 the table is settled at once, with no statistics read:
// its selectivity (which only orders fused predicates) is the estimate, and
// a pair table's layout needs no count.  Its 16-bit values are taken only
// when the group-by column's range fits them (vmax - vmin < 0xFFFF), so no
// value overflows; its set slots are at most the build rows (each inserts one
// key), so a dictionary sized for that many values holds them all.
static std::shared_ptr<CompactState> build_compact(Engine &eng, const void *ht, bool pairs, int32_t ht_min,
                                                   int32_t ht_max, int32_t vmin, int32_t vmax, uint64_t build_rows,
                                                   float est_sel = -1.0f)
{
    auto c = std::make_shared<CompactState>();
    const int64_t len = (int64_t)ht_max - ht_min + 1;
    const uint64_t words = (uint64_t)(len + 31) / 32;
    uint32_t *bits = eng.mm->alloc<uint32_t>((words + 3) / 4 * 4);
    // value overflow (cleared), then one count of set slots per block
    const uint32_t nparts = bits_partials((int32_t)len);
    uint32_t *stats = eng.mm->alloc_zero<uint32_t>(1 + (uint64_t)nparts);
    const bool narrow = pairs && (int64_t)vmax - vmin < 0xFFFF;
    uint16_t *vals = narrow ? eng.mm->alloc<uint16_t>((uint64_t)len) : nullptr;
    if (pairs)
        launch_bits_from_pairs(static_cast<const int32_t *>(ht), (int32_t)len, bits, vals, vmin, stats + 1, stats,
                               eng.stream);
    else
        launch_bits_from_keys(static_cast<const uint8_t *>(ht), (int32_t)len, bits, stats + 1, eng.stream);
    eng.launches++;
    c->bits = bits;
    c->stats = stats;
    c->nparts = nparts;
    c->pairs = pairs;
    c->ht = ht;
    c->narrow = vals;
    c->words = words;
    c->build_rows = build_rows;
    // the dictionary goes into a buffer of the largest size staged in LDS;
    // whether it is used depends on the number of set slots
    if (narrow && words <= 0xFFFFu && rank_values_off((uint32_t)words) < F_RANK_DICT_MAX) {
        c->dict = eng.mm->alloc<uint8_t>(F_RANK_DICT_MAX);
        c->dict_cap = F_RANK_DICT_MAX;
        launch_rank_dict(bits, (int32_t)len, vals, stats + 1, c->dict, c->dict_cap, eng.stream);
        eng.launches++;
    }
    if (est_sel >= 0.0f) {
        c->sel = std::min(1.0f, est_sel);
        if (pairs) settle_values(*c, (uint32_t)std::min<uint64_t>(build_rows, (uint64_t)len), 0);
        c->resolved = true;
    } else {
        eng.pending_compact.push_back(c);
    }
    return c;
}

void Engine::resolve_compact()
{
    std::vector<std::shared_ptr<CompactState>> todo;
    for (auto &w : pending_compact)
        if (auto c = w.lock())
            if (!c->resolved) todo.push_back(c);
    pending_compact.clear();
    if (todo.empty()) return;
    // each table's counts summed on the device (PACK_MAX_PARTS tables per
    // launch) into one buffer of (set slots, overflow) pairs: one transfer,
    // one sync
    uint32_t *staged = static_cast<uint32_t *>(pinned_buffer(8 * todo.size()));
    uint32_t *dev = mm->alloc<uint32_t>(2 * todo.size());
    StatsParts sp{};
    size_t base = 0;
    for (size_t i = 0; i < todo.size(); i++) {
        sp.stats[sp.n] = todo[i]->stats;
        sp.nparts[sp.n] = todo[i]->nparts;
        if (++sp.n == PACK_MAX_PARTS || i + 1 == todo.size()) {
            launch_stats_reduce(sp, dev + 2 * base, stream);
            launches++;
            base = i + 1;
            sp = StatsParts{};
        }
    }
    SDB_HIP(hipMemcpyAsync(staged, dev, 8 * todo.size(), hipMemcpyDeviceToHost, stream));
    mark("stats wait");
    wait_stream();
    mark("stats ready");
    const std::vector<uint32_t> host(staged, staged + 2 * todo.size());
    for (size_t i = 0; i < todo.size(); i++) {
        CompactState &c = *todo[i];
        const uint32_t set = host[2 * i], overflow = host[2 * i + 1];
        // fraction of the build rows selected: the share of uniformly drawn
        // foreign keys that find a match
        c.sel = std::min(1.0f, (float)set / (float)std::max<uint64_t>(c.build_rows, 1));
        if (c.pairs) settle_values(c, set, overflow);
        c.resolved = true;
    }
}

void TransientTable::apply_join(TransientTable &right, const RelNode &rel)
{
    if (rel.condition.operands.size() < 2) throw std::invalid_argument("join condition needs two operands");
    const int64_t left_column = rel.condition.operands[0].input;
    const int64_t right_column = rel.condition.operands[1].input - (int64_t)current_columns_.size();
    if (left_column < 0 || left_column >= (int64_t)current_columns_.size() || right_column < 0 ||
        right_column >= (int64_t)right.current_columns_.size())
        throw std::invalid_argument("Invalid column indices in join condition.");
    Column *probe = col(left_column, "join probe");

    if (rel.joinType == "semi") {
        auto ht = right.build_keys_hash_table((int)right_column);
        ensure_secondary_flags();
        right.execute_pending_kernels(false);
        std::shared_ptr<CompactState> ct;
        if (eng_.fusion)
            ct = build_compact(eng_, std::get<0>(ht), false, std::get<1>(ht), std::get<2>(ht), 0, 0, right.nrows_,
                               join_estimates() ? right.est_kept_ : -1.0f);
        std::vector<KernelBundle> ops(segment_num());
        for (uint64_t s = 0; s < segment_num(); s++) {
            const uint64_t off = s * eng_.segment_size;
            KernelData k;
            k.type = KernelType::FilterJoinKernel;
            k.n = (int64_t)seg_rows(s);
            k.a = probe->i32(s);
            k.in_flags = flags_ + off;
            k.out_flags = flags_secondary_ + off;
            k.ht = std::get<0>(ht);
            k.ht_min = std::get<1>(ht);
            k.ht_max = std::get<2>(ht);
            if (ct) k.bits = ct->bits;
            k.compact = ct;
            ops[s].add_kernel(k);
        }
        add_pending_kernels(std::move(ops));
        swap_flag_buffers();
        for (size_t i = 0; i < right.current_columns_.size(); i++) current_columns_.push_back(nullptr);
        return;
    }

    if (!right.group_by_column_) throw std::runtime_error("full join: the build table has no group-by column");
    Column &nc = new_column(nrows_);
    nc.data = eng_.row_buffer<int32_t>(name_, std::max<uint64_t>(nrows_, 1));
    const auto gb_mm = right.group_by_column_->get_min_max();
    const uint64_t gb_index = current_columns_.size() + right.group_by_column_index_;
    for (size_t i = 0; i < right.current_columns_.size(); i++) current_columns_.push_back(nullptr);
    if (gb_index >= current_columns_.size()) throw std::runtime_error("full join: group-by column index out of range");
    current_columns_[gb_index] = &nc;
    auto ht = right.build_key_vals_hash_table((int)right_column);
    ensure_secondary_flags();
    right.execute_pending_kernels(false);
    std::shared_ptr<CompactState> ct;
    if (eng_.fusion)
        ct = build_compact(eng_, std::get<0>(ht), true, std::get<1>(ht), std::get<2>(ht), gb_mm.first, gb_mm.second,
                           right.nrows_, join_estimates() ? right.est_kept_ : -1.0f);
    std::vector<KernelBundle> ops(segment_num());
    for (uint64_t s = 0; s < segment_num(); s++) {
        const uint64_t off = s * eng_.segment_size;
        KernelData k;
        k.type = KernelType::FullJoinKernel;
        k.n = (int64_t)seg_rows(s);
        k.a = probe->i32(s);
        k.result = nc.i32(s);
        k.in_flags = flags_ + off;
        k.out_flags = flags_secondary_ + off;
        k.ht = std::get<0>(ht);
        k.ht_min = std::get<1>(ht);
        k.ht_max = std::get<2>(ht);
        if (ct) k.bits = ct->bits;
        k.compact = ct;
        k.vmin = gb_mm.first;
        nc.seg_min[s] = gb_mm.first; // models/models.hpp:1609-1610
        nc.seg_max[s] = gb_mm.second;
        ops[s].add_kernel(k);
    }
    add_pending_kernels(std::move(ops));
    swap_flag_buffers();
}

} // namespace sdb
