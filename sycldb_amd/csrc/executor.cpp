// Plan analysis and the plan walk: operations/preprocessing.cpp:14-275 and
// executor/ddor_executor.cpp:22-245.
#include <cstring>
#include <deque>
#include <iostream>
#include <set>

#include "engine.hpp"

namespace sdb {

// kernels/types.cpp:3-21
const std::map<std::string, int> TABLE_COLUMN_NUMBERS = {
    {"lineorder", 17}, {"part", 9}, {"supplier", 7}, {"customer", 8}, {"ddate", 17}};
const std::map<std::string, std::vector<int>> TABLE_COLUMN_INDICES = {
    {"lineorder", {2, 3, 4, 5, 8, 9, 11, 12, 13}},
    {"part", {0, 2, 3, 4}},
    {"supplier", {0, 3, 4, 5}},
    {"customer", {0, 3, 4, 5}},
    {"ddate", {0, 4, 5}}};

namespace {

int column_count(const std::string &table)
{
    auto it = TABLE_COLUMN_NUMBERS.find(table);
    if (it == TABLE_COLUMN_NUMBERS.end()) throw std::invalid_argument("unknown table " + table);
    return it->second;
}

void expression_columns(const ExprType &e, std::set<int64_t> &out)
{
    if (e.exprType == ExprOption::COLUMN) out.insert(e.input);
    else if (e.exprType == ExprOption::EXPR)
        for (const ExprType &o : e.operands) expression_columns(o, out);
}

} // namespace

std::vector<int> dag_topological_sort(const PlanResult &plan)
{
    const size_t n = plan.rels.size();
    std::vector<std::vector<char>> dag(n, std::vector<char>(n, 0));
    std::deque<int> S;
    auto edge = [&](int64_t a, int64_t b) {
        if (a < 0 || b < 0 || a >= (int64_t)n || b >= (int64_t)n) throw std::invalid_argument("plan: bad rel input");
        dag[a][b] = 1;
    };
    for (const RelNode &rel : plan.rels) {
        switch (rel.relOp) {
        case RelNodeType::TABLE_SCAN: S.push_front((int)rel.id); break;
        case RelNodeType::FILTER:
        case RelNodeType::PROJECT:
        case RelNodeType::AGGREGATE:
        case RelNodeType::SORT: edge(rel.id - 1, rel.id); break;
        case RelNodeType::JOIN:
            if (rel.inputs.size() != 2) return {};
            edge(rel.inputs[0], rel.id);
            edge(rel.inputs[1], rel.id);
            break;
        default: return {};
        }
    }
    std::vector<int> sorted;
    while (!S.empty()) {
        const int node = S.back();
        S.pop_back();
        const RelNode &r = plan.rels[node];
        // the lineorder scan waits until it is the only source left (60-66)
        if (!S.empty() && r.relOp == RelNodeType::TABLE_SCAN && r.tables.size() > 1 && r.tables[1] == "lineorder") {
            S.push_front(node);
            continue;
        }
        sorted.push_back(node);
        for (size_t i = 0; i < n; i++) {
            if (!dag[node][i]) continue;
            dag[node][i] = 0;
            bool inbound = false;
            for (size_t j = 0; j < n && !inbound; j++) inbound = dag[j][i] != 0;
            if (!inbound) S.push_front((int)i);
        }
    }
    if (sorted.size() != n) return {};
    return sorted;
}

ExecutionInfo parse_execution_info(const PlanResult &plan)
{
    ExecutionInfo info;
    using Info = std::vector<std::pair<std::string, int>>;
    std::vector<Info> ops;
    auto at = [](const Info &v, int64_t i) -> const std::pair<std::string, int> & {
        if (i < 0 || i >= (int64_t)v.size()) throw std::invalid_argument("plan: column reference out of range");
        return v[i];
    };
    for (const RelNode &rel : plan.rels) {
        switch (rel.relOp) {
        case RelNodeType::TABLE_SCAN: {
            if (rel.tables.size() < 2) throw std::invalid_argument("plan: TABLE_SCAN needs [schema, table]");
            Info t;
            for (int i = 0; i < column_count(rel.tables[1]); i++) t.emplace_back(rel.tables[1], i);
            info.loaded_columns[rel.tables[1]];
            ops.push_back(t);
            break;
        }
        case RelNodeType::FILTER: {
            if (ops.empty()) throw std::invalid_argument("plan: FILTER without input");
            Info t = ops.back();
            std::set<int64_t> cols;
            expression_columns(rel.condition, cols);
            for (int64_t c : cols) info.loaded_columns[at(t, c).first].push_back(at(t, c).second);
            ops.push_back(t);
            break;
        }
        case RelNodeType::PROJECT: {
            if (ops.empty()) throw std::invalid_argument("plan: PROJECT without input");
            const Info last = ops.back();
            Info t;
            for (const ExprType &e : rel.exprs) {
                std::set<int64_t> cols;
                expression_columns(e, cols);
                for (int64_t c : cols) info.loaded_columns[at(last, c).first].push_back(at(last, c).second);
                if (!cols.empty()) t.push_back(at(last, *cols.begin()));
            }
            ops.push_back(t);
            break;
        }
        case RelNodeType::AGGREGATE: {
            if (ops.empty()) throw std::invalid_argument("plan: AGGREGATE without input");
            const Info last = ops.back();
            Info t;
            for (int64_t g : rel.group) {
                const auto &tc = at(last, g);
                info.loaded_columns[tc.first].push_back(tc.second);
                t.push_back(tc);
                info.group_by_columns[tc.first] = tc.second;
            }
            for (const AggType &a : rel.aggs) {
                for (int64_t c : a.operands) info.loaded_columns[at(last, c).first].push_back(at(last, c).second);
                if (!a.operands.empty()) t.push_back(at(last, a.operands[0]));
            }
            ops.push_back(t);
            break;
        }
        case RelNodeType::JOIN: {
            if (rel.inputs.size() != 2) throw std::invalid_argument("plan: JOIN needs two inputs");
            for (int64_t in : rel.inputs)
                if (in < 0 || in >= (int64_t)ops.size()) throw std::invalid_argument("plan: JOIN input out of range");
            Info t = ops[rel.inputs[0]];
            const Info &r = ops[rel.inputs[1]];
            t.insert(t.end(), r.begin(), r.end());
            ops.push_back(t);
            break;
        }
        case RelNodeType::SORT: {
            if (ops.empty()) throw std::invalid_argument("plan: SORT without input");
            ops.push_back(ops.back());
            break;
        }
        default: std::cerr << "Unsupported RelNodeType" << std::endl; break;
        }
    }
    info.dag_order = dag_topological_sort(plan);
    return info;
}

FinalTable run_ddor_plan(Engine &eng, const PlanResult &plan)
{
    const ExecutionInfo info = parse_execution_info(plan);
    if (info.dag_order.empty()) throw std::runtime_error("DAG topological sort failed");
    std::vector<int> output_table(plan.rels.size(), -1);
    std::deque<TransientTable> tts;
    for (const RelNode &rel : plan.rels) {
        if (rel.relOp != RelNodeType::TABLE_SCAN) continue;
        Table *t = eng.find_table(rel.tables[1]);
        if (!t) throw std::invalid_argument("Table " + rel.tables[1] + " not found among loaded tables.");
        TransientTable &tt = tts.emplace_back(eng, t);
        auto gb = info.group_by_columns.find(rel.tables[1]);
        if (gb != info.group_by_columns.end()) tt.set_group_by_column(gb->second);
        output_table[rel.id] = (int)tts.size() - 1;
    }
    auto input_of = [&](int64_t id) -> TransientTable & {
        if (id < 0 || id >= (int64_t)output_table.size() || output_table[id] < 0)
            throw std::invalid_argument("plan: operator input has no table");
        return tts[output_table[id]];
    };
    for (int id : info.dag_order) {
        const RelNode &rel = plan.rels[id];
        switch (rel.relOp) {
        case RelNodeType::TABLE_SCAN: break;
        case RelNodeType::FILTER:
            input_of(id - 1).apply_filter(rel.condition, "");
            output_table[id] = output_table[id - 1];
            break;
        case RelNodeType::PROJECT:
            input_of(id - 1).apply_project(rel.exprs);
            output_table[id] = output_table[id - 1];
            break;
        case RelNodeType::AGGREGATE:
            if (rel.aggs.empty()) throw std::invalid_argument("plan: AGGREGATE without aggs");
            input_of(id - 1).apply_aggregate(rel.aggs[0], rel.group);
            output_table[id] = output_table[id - 1];
            break;
        case RelNodeType::JOIN:
            input_of(rel.inputs[0]).apply_join(input_of(rel.inputs[1]), rel);
            output_table[id] = output_table[rel.inputs[0]];
            break;
        case RelNodeType::SORT: output_table[id] = output_table[id - 1]; break;
        default: std::cerr << "RelNodeType not yet supported in DDOR." << std::endl; break;
        }
    }
    TransientTable &final_table = input_of((int64_t)plan.rels.size() - 1);
    final_table.execute_pending_kernels(false);

    FinalTable t;
    t.nrows = final_table.get_nrows();
    t.flags = final_table.flags_device();
    for (Column *c : final_table.get_columns()) {
        if (!c || !c->loaded || !c->data) continue;
        t.cols.push_back({c->data, c->is_aggregate, std::min(t.nrows, c->nrows)});
    }
    return t;
}

void Engine::print_marks()
{
    if (marks.empty()) return;
    const auto t0 = marks.front().first;
    std::string line = "[step]";
    for (const auto &m : marks) {
        char buf[64];
        snprintf(buf, sizeof(buf), " %.1f", std::chrono::duration<double, std::micro>(m.first - t0).count());
        line += buf;
        line += " " + m.second + ";";
    }
    fprintf(stderr, "%s\n", line.c_str());
    marks.clear();
}

void Engine::wait_stream()
{
    if (!sync_ev) SDB_HIP(hipEventCreateWithFlags(&sync_ev, hipEventDisableTiming));
    hipEvent_t ev = stop_at_wait ? ev_stop : sync_ev;
    SDB_HIP(hipEventRecord(ev, stream));
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    SDB_HIP(e);
    if (stop_at_wait) stop_recorded = true;
}

void *Engine::pinned_buffer(size_t bytes)
{
    if (bytes > pinned_bytes) {
        if (pinned) SDB_HIP(hipHostFree(pinned));
        pinned = nullptr;
        pinned_bytes = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 20);
        SDB_HIP(hipHostMalloc(&pinned, want, hipHostMallocDefault));
        pinned_bytes = want;
    }
    return pinned;
}

namespace {

// Group-by results of at least this many slots come back compacted on the
// device: only their selected rows cross PCIe and reach host memory.
constexpr uint64_t RESULT_SPARSE_MIN_ROWS = 1u << 15;

bool sparse_result(const FinalTable &t)
{
    if (t.nrows < RESULT_SPARSE_MIN_ROWS || t.nrows >= (1ull << 31) || !t.flags ||
        t.cols.size() > (size_t)GATHER_MAX_COLS)
        return false;
    for (const FinalTable::Col &c : t.cols)
        if (c.is_u64) return true;
    return false;
}

// Device-to-host copies through the engine's pinned staging buffer: every
// copy queued, one wait for the stream, then into host memory.  Up to
// PACK_MAX_BYTES in all, the arrays are first packed on the device.
constexpr size_t PACK_MAX_BYTES = size_t(64) << 20;
struct HostCopies {
    struct Part {
        const void *src;
        void *dst;
        size_t off, bytes;
    };
    std::vector<Part> parts;
    size_t total = 0;
    void add(const void *src, void *dst, size_t bytes)
    {
        if (!bytes) return;
        total = (total + 15) / 16 * 16;
        parts.push_back({src, dst, total, bytes});
        total += bytes;
    }
    void run(Engine &eng)
    {
        if (parts.empty()) {
            eng.wait_stream();
            return;
        }
        uint8_t *host = static_cast<uint8_t *>(eng.pinned_buffer(total));
        if (parts.size() == 1 || total > PACK_MAX_BYTES) {
            for (const Part &p : parts)
                SDB_HIP(hipMemcpyAsync(host + p.off, p.src, p.bytes, hipMemcpyDeviceToHost, eng.stream));
        } else {
            // several arrays: packed on the device into the staging layout,
            // then one transfer (each transfer costs ~10 us of latency)
            uint8_t *dev = eng.mm->alloc<uint8_t>(total);
            PackParts pk{};
            for (const Part &p : parts) {
                pk.src[pk.n] = static_cast<const uint8_t *>(p.src);
                pk.off[pk.n] = p.off;
                pk.bytes[pk.n] = p.bytes;
                if (++pk.n == PACK_MAX_PARTS) {
                    launch_pack_parts(pk, dev, eng.stream);
                    eng.launches++;
                    pk.n = 0;
                }
            }
            if (pk.n) {
                launch_pack_parts(pk, dev, eng.stream);
                eng.launches++;
            }
            SDB_HIP(hipMemcpyAsync(host, dev, total, hipMemcpyDeviceToHost, eng.stream));
        }
        eng.mark("result wait");
        eng.wait_stream();
        eng.mark("result ready");
        for (const Part &p : parts) std::memcpy(p.dst, host + p.off, p.bytes);
        parts.clear();
        total = 0;
    }
};

} // namespace

std::vector<ResultTable> copy_results(Engine &eng, const std::vector<FinalTable> &ts)
{
    // host copies of the final tables (the materialize_host_view +
    // save_result path, executor/result_writer.cpp:28-67), all tables in
    // at most two waits for the stream
    std::vector<ResultTable> out(ts.size());
    HostCopies copies;
    std::vector<size_t> sparse;
    for (size_t i = 0; i < ts.size(); i++)
        if (sparse_result(ts[i])) sparse.push_back(i);
    // large group-by tables: the selected row ids, compacted in row order on
    // the device; their counts come back with the dense tables
    std::vector<int32_t *> ids(ts.size(), nullptr);
    std::vector<uint32_t> counts(sparse.size(), 0);
    if (!sparse.empty()) {
        uint32_t *d_counts = eng.mm->alloc<uint32_t>(sparse.size());
        for (size_t j = 0; j < sparse.size(); j++) {
            const FinalTable &t = ts[sparse[j]];
            eng.registry.check(t.flags, t.nrows, "result flags");
            int32_t *row_ids = eng.mm->alloc<int32_t>(t.nrows);
            uint32_t *tiles = eng.mm->alloc<uint32_t>((uint64_t)std::max<int64_t>(compact_tiles((int64_t)t.nrows), 1));
            launch_compact_into(t.flags, (int64_t)t.nrows, row_ids, d_counts + j, tiles, eng.stream);
            eng.launches += 3;
            ids[sparse[j]] = row_ids;
        }
        copies.add(d_counts, counts.data(), counts.size() * 4);
    }
    for (size_t i = 0; i < ts.size(); i++) {
        const FinalTable &t = ts[i];
        ResultTable &res = out[i];
        res.nrows = t.nrows;
        res.sparse = ids[i] != nullptr;
        for (const FinalTable::Col &c : t.cols) {
            ResultTable::Col rc;
            rc.is_u64 = c.is_u64;
            res.cols.push_back(std::move(rc));
        }
        if (res.sparse) continue;
        res.flags.resize(res.nrows);
        if (res.nrows) {
            eng.registry.check(t.flags, res.nrows, "result flags");
            copies.add(t.flags, res.flags.data(), res.nrows);
        }
        for (size_t c = 0; c < t.cols.size(); c++) {
            const FinalTable::Col &fc = t.cols[c];
            ResultTable::Col &rc = res.cols[c];
            const uint64_t elem = rc.is_u64 ? 8 : 4;
            if (rc.is_u64) rc.u64.assign(res.nrows, 0);
            else rc.i32.assign(res.nrows, 0);
            if (fc.rows) {
                eng.registry.check(fc.data, fc.rows * elem, "result column");
                copies.add(fc.data, rc.is_u64 ? (void *)rc.u64.data() : (void *)rc.i32.data(), fc.rows * elem);
            }
        }
    }
    copies.run(eng);
    // large group-by tables: every column gathered at the selected rows on
    // the device, then copied
    bool any = false;
    for (size_t j = 0; j < sparse.size(); j++) {
        const FinalTable &t = ts[sparse[j]];
        ResultTable &res = out[sparse[j]];
        const uint32_t k = counts[j];
        if (k > t.nrows) fail(SDB_ERR_RUNTIME, "result compaction: selected count out of range");
        res.sel.resize(k);
        if (!k) continue;
        any = true;
        GatherCols g{};
        g.ncols = (int32_t)t.cols.size();
        for (size_t c = 0; c < t.cols.size(); c++) {
            const FinalTable::Col &fc = t.cols[c];
            ResultTable::Col &rc = res.cols[c];
            const uint64_t elem = fc.is_u64 ? 8 : 4;
            if (fc.rows) eng.registry.check(fc.data, fc.rows * elem, "result column");
            g.src[c] = fc.data;
            g.src_rows[c] = fc.rows;
            if (fc.is_u64) g.is_u64 |= 1u << c;
            void *d = eng.mm->alloc<uint8_t>((uint64_t)k * elem);
            g.dst[c] = d;
            if (fc.is_u64) rc.u64.resize(k);
            else rc.i32.resize(k);
            copies.add(d, fc.is_u64 ? (void *)rc.u64.data() : (void *)rc.i32.data(), (size_t)k * elem);
        }
        launch_gather_cols(g, ids[sparse[j]], (int64_t)k, eng.stream);
        eng.launches++;
        copies.add(ids[sparse[j]], res.sel.data(), (size_t)k * 4);
    }
    if (any) copies.run(eng);
    return out;
}

ResultTable copy_result(Engine &eng, const FinalTable &t) { return std::move(copy_results(eng, {t})[0]); }

uint64_t ResultTable::selected() const
{
    if (sparse) return sel.size();
    uint64_t n = 0;
    for (uint8_t f : flags) n += f != 0;
    return n;
}

void ResultTable::flags_to(uint8_t *host) const
{
    if (!sparse) {
        std::memcpy(host, flags.data(), flags.size());
        return;
    }
    std::memset(host, 0, nrows);
    for (uint32_t r : sel) host[r] = 1;
}

void ResultTable::col_to(int c, void *host) const
{
    const Col &col = cols[c];
    if (!sparse) {
        if (col.is_u64) std::memcpy(host, col.u64.data(), col.u64.size() * 8);
        else std::memcpy(host, col.i32.data(), col.i32.size() * 4);
        return;
    }
    if (col.is_u64) {
        uint64_t *h = static_cast<uint64_t *>(host);
        std::memset(h, 0, nrows * 8);
        for (size_t j = 0; j < sel.size(); j++) h[sel[j]] = col.u64[j];
    } else {
        int32_t *h = static_cast<int32_t *>(host);
        std::memset(h, 0, nrows * 4);
        for (size_t j = 0; j < sel.size(); j++) h[sel[j]] = col.i32[j];
    }
}

void ResultTable::selected_rows_to(uint64_t *host) const
{
    if (sparse) {
        for (size_t j = 0; j < sel.size(); j++) host[j] = sel[j];
        return;
    }
    uint64_t j = 0;
    for (uint64_t r = 0; r < flags.size(); r++)
        if (flags[r]) host[j++] = r;
}

void ResultTable::col_selected_to(int c, void *host) const
{
    const Col &col = cols[c];
    if (sparse) {
        if (col.is_u64) std::memcpy(host, col.u64.data(), col.u64.size() * 8);
        else std::memcpy(host, col.i32.data(), col.i32.size() * 4);
        return;
    }
    uint64_t j = 0;
    for (uint64_t r = 0; r < flags.size(); r++) {
        if (!flags[r]) continue;
        if (col.is_u64) static_cast<uint64_t *>(host)[j++] = col.u64[r];
        else static_cast<int32_t *>(host)[j++] = col.i32[r];
    }
}

ResultTable execute_ddor_plan(Engine &eng, const PlanResult &plan) { return copy_result(eng, run_ddor_plan(eng, plan)); }

} // namespace sdb
