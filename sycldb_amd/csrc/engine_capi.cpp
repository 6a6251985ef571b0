// C ABI of the engine: resident tables, plan execution, results.
#include <chrono>
#include <cstring>
#include <cstdio>
#include <fstream>

#include "capi_util.hpp"
#include "engine.hpp"
#include "fused.hpp"

#include "capi_handles.hpp"

using namespace sdb;

namespace sdb {

Table *Engine::find_table(const std::string &name)
{
    auto it = tables.find(name);
    return it == tables.end() ? nullptr : it->second.get();
}

void Engine::prof_begin(const std::string &name, int64_t rows, ProfRec &rec)
{
    auto take = [&]() {
        hipEvent_t e;
        if (!event_pool.empty()) {
            e = event_pool.back();
            event_pool.pop_back();
        } else {
            SDB_HIP(hipEventCreate(&e));
        }
        return e;
    };
    rec.name = name;
    rec.rows = rows;
    rec.a = take();
    rec.b = take();
    SDB_HIP(hipEventRecord(rec.a, stream));
}

void Engine::prof_end(ProfRec &rec)
{
    SDB_HIP(hipEventRecord(rec.b, stream));
    prof_pending.push_back(rec);
}

void Engine::prof_collect()
{
    for (ProfRec &r : prof_pending) {
        // polled, as the engine's other waits: a blocking wait's wake-up
        // cost tens of microseconds per step even for a completed event
        hipError_t q;
        while ((q = hipEventQuery(r.b)) == hipErrorNotReady) {
        }
        SDB_HIP(q);
        float ms = 0;
        SDB_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        KernelStat &st = prof_stats[r.name];
        st.launches++;
        st.rows += r.rows;
        st.ms += ms;
        event_pool.push_back(r.a);
        event_pool.push_back(r.b);
    }
    prof_pending.clear();
}

} // namespace sdb

namespace {

Table &table_of(sdb_engine *e, const char *name)
{
    require(e && name, "null engine or name");
    Table *t = e->eng.find_table(name);
    if (!t) fail(SDB_ERR_INVALID_ARG, std::string("unknown table ") + name);
    return *t;
}

Column &column_of(sdb_engine *e, const char *name, int col)
{
    Table &t = table_of(e, name);
    require(col >= 0 && col < (int)t.columns.size(), "column index out of range");
    return t.columns[col];
}

void ensure_storage(Engine &eng, Table &t, Column &c)
{
    if (c.data) return;
    void *p = nullptr;
    SDB_HIP(hipMalloc(&p, std::max<uint64_t>(t.nrows, 1) * sizeof(int32_t)));
    eng.registry.add(p, std::max<uint64_t>(t.nrows, 1) * sizeof(int32_t));
    t.owned.push_back(p);
    c.data = p;
    c.loaded = true;
    c.nrows = t.nrows;
    c.segment_size = eng.segment_size;
}

} // namespace

extern "C" {

int sdb_engine_create(const sdb_engine_options *opts, sdb_engine **out)
{
    return guarded([&] {
        require(out, "null output");
        *out = nullptr;
        auto e = std::make_unique<sdb_engine>();
        Engine &eng = e->eng;
        uint64_t budget = 16ull << 30;
        if (opts) {
            if (opts->device_budget_bytes) budget = opts->device_budget_bytes;
            if (opts->segment_size) eng.segment_size = opts->segment_size;
            eng.fusion = opts->disable_fusion == 0;
            eng.device = opts->device;
        }
        require(eng.segment_size >= 1 && eng.segment_size <= (1ull << 31), "segment size must be in [1, 2^31]");
        SDB_HIP(hipSetDevice(eng.device));
        SDB_HIP(hipStreamCreateWithFlags(&eng.stream, hipStreamNonBlocking));
        SDB_HIP(hipEventCreate(&eng.ev_start));
        SDB_HIP(hipEventCreate(&eng.ev_stop));
        eng.mm = std::make_unique<MemoryManager>(budget, eng.stream, &eng.registry);
        *out = e.release();
    });
}

void sdb_engine_destroy(sdb_engine *e)
{
    if (!e) return;
    Engine &eng = e->eng;
    (void)hipSetDevice(eng.device);
    if (eng.stream) (void)hipStreamSynchronize(eng.stream);
    eng.mm.reset();
    for (auto &kv : eng.tables)
        for (void *p : kv.second->owned) (void)hipFree(p);
    for (auto &r : eng.prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t ev : eng.event_pool) (void)hipEventDestroy(ev);
    if (eng.pinned) (void)hipHostFree(eng.pinned);
    if (eng.sync_ev) (void)hipEventDestroy(eng.sync_ev);
    if (eng.ev_start) (void)hipEventDestroy(eng.ev_start);
    if (eng.ev_stop) (void)hipEventDestroy(eng.ev_stop);
    if (eng.stream) (void)hipStreamDestroy(eng.stream);
    delete e;
}

sdb_stream sdb_engine_stream(sdb_engine *e) { return e ? reinterpret_cast<sdb_stream>(e->eng.stream) : nullptr; }

int sdb_engine_table_create(sdb_engine *e, const char *name, uint64_t nrows)
{
    return guarded([&] {
        require(e && name, "null engine or name");
        auto it = TABLE_COLUMN_NUMBERS.find(name);
        if (it == TABLE_COLUMN_NUMBERS.end()) fail(SDB_ERR_INVALID_ARG, std::string("unknown table ") + name);
        Engine &eng = e->eng;
        SDB_HIP(hipSetDevice(eng.device));
        auto old = eng.tables.find(name);
        if (old != eng.tables.end()) {
            SDB_HIP(hipStreamSynchronize(eng.stream));
            for (void *p : old->second->owned) {
                eng.registry.remove(p);
                SDB_HIP(hipFree(p));
            }
            eng.tables.erase(old);
        }
        auto t = std::make_unique<Table>();
        t->name = name;
        t->nrows = nrows;
        t->columns.resize(it->second);
        for (Column &c : t->columns) {
            c.nrows = nrows;
            c.segment_size = eng.segment_size;
        }
        eng.tables[name] = std::move(t);
    });
}

int sdb_engine_column_alloc(sdb_engine *e, const char *name, int col, int32_t **dev_ptr)
{
    return guarded([&] {
        require(dev_ptr, "null output");
        Table &t = table_of(e, name);
        Column &c = column_of(e, name, col);
        SDB_HIP(hipSetDevice(e->eng.device));
        ensure_storage(e->eng, t, c);
        *dev_ptr = c.i32();
    });
}

int sdb_engine_column_upload(sdb_engine *e, const char *name, int col, const int32_t *host)
{
    return guarded([&] {
        require(host, "null host data");
        Table &t = table_of(e, name);
        Column &c = column_of(e, name, col);
        SDB_HIP(hipSetDevice(e->eng.device));
        ensure_storage(e->eng, t, c);
        if (t.nrows)
            SDB_HIP(hipMemcpyAsync(c.data, host, t.nrows * 4, hipMemcpyHostToDevice, e->eng.stream));
        SDB_HIP(hipStreamSynchronize(e->eng.stream));
    });
}

int sdb_engine_column_download(sdb_engine *e, const char *name, int col, int32_t *host)
{
    return guarded([&] {
        require(host, "null host buffer");
        Table &t = table_of(e, name);
        Column &c = column_of(e, name, col);
        require(c.data != nullptr, "column has no storage");
        SDB_HIP(hipSetDevice(e->eng.device));
        if (t.nrows)
            SDB_HIP(hipMemcpyAsync(host, c.data, t.nrows * 4, hipMemcpyDeviceToHost, e->eng.stream));
        SDB_HIP(hipStreamSynchronize(e->eng.stream));
    });
}

int sdb_engine_table_load_dir(sdb_engine *e, const char *name, const char *dir)
{
    return guarded([&] {
        require(e && name && dir, "null argument");
        auto idx = TABLE_COLUMN_INDICES.find(name);
        if (idx == TABLE_COLUMN_INDICES.end()) fail(SDB_ERR_INVALID_ARG, std::string("unknown table ") + name);
        // models/models.hpp:1659-1708: <DIR>/<TABLE><col>, raw int32, for
        // the columns of table_column_indices in ascending order.  A file
        // that cannot be read counts as 0 entries.  The first column's entry
        // count fixes the row count; a column of another count stays
        // unloaded, with a warning.
        std::string upper = name;
        for (char &ch : upper) ch = (char)toupper((unsigned char)ch);
        std::vector<std::pair<int, std::vector<int32_t>>> cols;
        int64_t nrows = -1;
        for (int ci : idx->second) {
            const std::string path = std::string(dir) + "/" + upper + std::to_string(ci);
            std::ifstream f(path, std::ios::binary | std::ios::ate);
            const std::streampos size = f ? f.tellg() : std::streampos(-1);
            const int64_t n = size == std::streampos(-1) ? 0 : (int64_t)size / 4;
            if (nrows < 0) nrows = n;
            if (n != nrows) {
                fprintf(stderr, "Warning: Column length mismatch in %s: expected %lld, got %lld\n", path.c_str(),
                        (long long)nrows, (long long)n);
                continue;
            }
            std::vector<int32_t> data((size_t)n);
            if (n > 0) {
                f.seekg(0);
                f.read(reinterpret_cast<char *>(data.data()), n * 4);
                if (!f) fail(SDB_ERR_RUNTIME, "reading column file " + path + " failed");
            }
            cols.emplace_back(ci, std::move(data));
        }
        if (nrows < 0) nrows = 0;
        int rc = sdb_engine_table_create(e, name, (uint64_t)nrows);
        if (rc) fail(rc, sdb_last_error());
        static const int32_t none = 0; // the source of an empty column
        for (auto &c : cols) {
            rc = sdb_engine_column_upload(e, name, c.first, c.second.empty() ? &none : c.second.data());
            if (rc) fail(rc, sdb_last_error());
        }
        rc = sdb_engine_table_finalize(e, name);
        if (rc) fail(rc, sdb_last_error());
    });
}

int sdb_engine_table_finalize(sdb_engine *e, const char *name)
{
    return guarded([&] {
        Table &t = table_of(e, name);
        Engine &eng = e->eng;
        SDB_HIP(hipSetDevice(eng.device));
        for (Column &c : t.columns) {
            if (!c.loaded || !c.data) continue;
            const uint64_t ns = c.nseg();
            c.seg_min.assign(ns, 0);
            c.seg_max.assign(ns, 0);
            for (uint64_t s = 0; s < ns; s++)
                minmax_i32(c.i32(s), (int64_t)c.seg_rows(s), &c.seg_min[s], &c.seg_max[s], eng.stream);
        }
    });
}

int sdb_engine_execute(sdb_engine *e, const char *plan_json, sdb_result **out)
{
    return guarded([&] {
        require(e && plan_json && out, "null argument");
        *out = nullptr;
        Engine &eng = e->eng;
        static const bool trace = [] {
            const char *v = getenv("SDB_TRACE_STEPS");
            return v && *v == '1';
        }();
        eng.tracing = trace;
        eng.mark("start");
        SDB_HIP(hipSetDevice(eng.device));
        const PlanResult plan = parse_plan(plan_json);
        eng.mark("parsed");
        // scratch of the previous query is released (runtime/runtime.cpp:331-345)
        eng.mm->reset(eng.stream);
        eng.launches = 0;
        const auto w0 = std::chrono::steady_clock::now();
        SDB_HIP(hipEventRecord(eng.ev_start, eng.stream));
        auto r = std::make_unique<sdb_result>();
        eng.stop_at_wait = true;
        eng.stop_recorded = false;
        try {
            r->t = execute_ddor_plan(eng, plan);
        } catch (...) {
            eng.stop_at_wait = false;
            throw;
        }
        eng.stop_at_wait = false;
        eng.mark("plan done");
        // the result wait recorded ev_stop; a stream with work since (still
        // running) gets its end recorded now
        const hipError_t idle = eng.stop_recorded ? hipStreamQuery(eng.stream) : hipErrorNotReady;
        if (idle != hipSuccess) {
            if (idle != hipErrorNotReady) SDB_HIP(idle);
            SDB_HIP(hipEventRecord(eng.ev_stop, eng.stream));
            hipError_t q;
            while ((q = hipEventQuery(eng.ev_stop)) == hipErrorNotReady) {
            }
            SDB_HIP(q);
        }
        if (eng.profile) eng.prof_collect();
        eng.mark("end");
        eng.print_marks();
        const auto w1 = std::chrono::steady_clock::now();
        float ms = 0;
        SDB_HIP(hipEventElapsedTime(&ms, eng.ev_start, eng.ev_stop));
        eng.last_device_ms = ms;
        eng.last_wall_ms = std::chrono::duration<double, std::milli>(w1 - w0).count();
        *out = r.release();
    });
}

} // extern "C"

// SDB_DEBUG_BATCH=1 reports on stderr why a plan runs outside the shared scan.
static bool debug_batch()
{
    static const bool on = [] {
        const char *e = getenv("SDB_DEBUG_BATCH");
        return e && *e == '1';
    }();
    return on;
}

// SDB_BATCH_MASKS=0 keeps every probe inside its program in batched launches
// (no shared probe masks).
static bool batch_masks()
{
    const char *e = getenv("SDB_BATCH_MASKS");
    return !(e && *e == '0');
}

namespace sdb {
static void *batch_alloc(void *ctx, size_t bytes) { return static_cast<Engine *>(ctx)->mm->alloc<uint8_t>(bytes); }
} // namespace sdb

// Batched execution: every plan's dimension work runs as usual; the
// fact-table bundles of the plans that fuse are captured and run together in
// one kernel per segment (one scan of the union of their fact columns); the
// rest run through the plain path.  Results are those of running each plan
// alone.
namespace sdb {
std::vector<ResultTable> execute_batch(Engine &eng, const std::vector<PlanResult> &parsed, int *batched)
{
    if (parsed.size() > F_BATCH_MAX) {
        // in groups of F_BATCH_MAX (each program's LDS share is fixed)
        std::vector<ResultTable> all;
        int total = 0;
        for (size_t k = 0; k < parsed.size(); k += F_BATCH_MAX) {
            std::vector<PlanResult> part(parsed.begin() + k, parsed.begin() + std::min(parsed.size(), k + F_BATCH_MAX));
            int b = 0;
            std::vector<ResultTable> r = execute_batch(eng, part, &b);
            total += b;
            for (auto &t : r) all.push_back(std::move(t));
        }
        if (batched) *batched = total;
        return all;
    }
    const int n = (int)parsed.size();
    if (batched) *batched = 0;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    eng.mm->reset(eng.stream);
    struct Pipe {
        FinalTable fin;
        std::vector<CapturedBundle> bundles;
        std::vector<std::function<void()>> deferred;
        bool batched = false;
    };
    std::vector<Pipe> pipes(n);
    const bool can = eng.fusion && eng.codegen != SDB_CODEGEN_INTERPRET && jit_available() && n > 1;
    const std::string fact = "lineorder";
    const uint32_t lds_each = F_BATCH_LDS_EACH;
    for (int i = 0; i < n && can; i++) {
        eng.capture = Capture{};
        eng.capture.active = true;
        eng.capture.table = fact;
        eng.capture.lds_budget = lds_each;
        try {
            FinalTable fin = run_ddor_plan(eng, parsed[i]);
            // one bundle per segment, in segment order
            bool ok = !eng.capture.bundles.empty();
            for (size_t b = 0; b < eng.capture.bundles.size() && ok; b++) ok = eng.capture.bundles[b].segment == b;
            if (ok) {
                pipes[i].fin = fin;
                pipes[i].bundles = std::move(eng.capture.bundles);
                pipes[i].deferred = std::move(eng.capture.deferred);
                pipes[i].batched = true;
            } else if (debug_batch()) {
                fprintf(stderr, "[batch] plan %d runs alone: %zu captured bundles\n", i, eng.capture.bundles.size());
            }
        } catch (const NotBatchable &ex) {
            if (debug_batch()) fprintf(stderr, "[batch] plan %d runs alone: %s\n", i, ex.what());
        } catch (...) {
            eng.capture = Capture{};
            throw;
        }
        eng.capture = Capture{};
    }
    // the captured programs, built now: every plan's join statistics are
    // read in one sync (resolve_compact) rather than one per plan
    if (can) {
        eng.capture = Capture{};
        eng.capture.active = true;
        eng.capture.table = fact;
        eng.capture.lds_budget = lds_each;
        try {
            for (int i = 0; i < n; i++) {
                if (!pipes[i].batched) continue;
                try {
                    for (CapturedBundle &cb : pipes[i].bundles) KernelBundle::build_captured(eng, cb);
                } catch (const NotBatchable &ex) {
                    if (debug_batch()) fprintf(stderr, "[batch] plan %d runs alone: %s\n", i, ex.what());
                    pipes[i].batched = false;
                }
            }
        } catch (...) {
            eng.capture = Capture{};
            throw;
        }
        eng.capture = Capture{};
    }
    const double t_capture = ms_since(t0);
    // the shared fact passes
    std::vector<int> members;
    for (int i = 0; i < n; i++)
        if (pipes[i].batched) members.push_back(i);
    if (!members.empty()) {
        const size_t nseg = pipes[members[0]].bundles.size();
        for (int i : members)
            if (pipes[i].bundles.size() != nseg) fail(SDB_ERR_RUNTIME, "batched pipelines disagree on segments");
        try {
        for (size_t sgi = 0; sgi < nseg; sgi++) {
            std::vector<FProgram> progs;
            const int64_t rows = pipes[members[0]].bundles[sgi].n;
            for (int i : members) {
                if (pipes[i].bundles[sgi].n != rows) fail(SDB_ERR_RUNTIME, "batched pipelines disagree on rows");
                progs.push_back(pipes[i].bundles[sgi].p);
            }
            Engine::ProfRec rec;
            const bool timed = eng.profiles(&fact);
            if (timed) eng.prof_begin("batch:" + fact, rows, rec);
            // one 4-row chunk per thread whatever the plain path's unroll: the
            // shared scan's double-buffered columns fill the registers
            launch_fused_batch(progs, rows, eng.stream, 1, eng.find_table(fact), batch_alloc, &eng,
                               batch_masks());
            if (timed) eng.prof_end(rec);
            eng.launches++;
        }
        } catch (const Error &e) {
            // a combination the batched kernel does not take (found before
            // its first launch): every member runs through the plain path
            if (e.code != SDB_ERR_UNSUPPORTED) throw;
            if (debug_batch()) fprintf(stderr, "[batch] shared pass refused: %s\n", e.what());
            for (int i : members) pipes[i].batched = false;
            members.clear();
        }
        for (int i : members)
            for (auto &f : pipes[i].deferred) f();
    }
    const double t_shared = ms_since(t0);
    for (int i = 0; i < n; i++)
        if (!pipes[i].batched) pipes[i].fin = run_ddor_plan(eng, parsed[i]);
    const double t_plain = ms_since(t0);
    std::vector<FinalTable> fins;
    for (int i = 0; i < n; i++) fins.push_back(pipes[i].fin);
    std::vector<ResultTable> out = copy_results(eng, fins);
    if (batched) *batched = (int)members.size();
    if (debug_batch())
        fprintf(stderr, "[batch] %d plans, %zu shared: host ms capture %.2f, shared pass %.2f, plain %.2f, results %.2f\n",
                n, members.size(), t_capture, t_shared - t_capture, t_plain - t_shared, ms_since(t0) - t_plain);
    return out;
}
} // namespace sdb

extern "C" {

int sdb_engine_execute_batch(sdb_engine *e, const char *const *plans, int n, sdb_result **out, int *batched)
{
    return guarded([&] {
        require(e && (n == 0 || (plans && out)) && n >= 0, "null argument");
        for (int i = 0; i < n; i++) out[i] = nullptr;
        if (batched) *batched = 0;
        Engine &eng = e->eng;
        SDB_HIP(hipSetDevice(eng.device));
        std::vector<PlanResult> parsed;
        for (int i = 0; i < n; i++) {
            require(plans[i] != nullptr, "null plan");
            parsed.push_back(parse_plan(plans[i]));
        }
        eng.launches = 0;
        const auto w0 = std::chrono::steady_clock::now();
        SDB_HIP(hipEventRecord(eng.ev_start, eng.stream));
        std::vector<ResultTable> res = execute_batch(eng, parsed, batched);
        SDB_HIP(hipEventRecord(eng.ev_stop, eng.stream));
        SDB_HIP(hipEventSynchronize(eng.ev_stop));
        if (eng.profile) eng.prof_collect();
        float ms = 0;
        SDB_HIP(hipEventElapsedTime(&ms, eng.ev_start, eng.ev_stop));
        eng.last_device_ms = ms;
        eng.last_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        for (int i = 0; i < n; i++) {
            auto r = std::make_unique<sdb_result>();
            r->t = std::move(res[i]);
            out[i] = r.release();
        }
    });
}

int sdb_engine_last_timing(sdb_engine *e, double *device_ms, double *wall_ms)
{
    return guarded([&] {
        require(e, "null engine");
        if (device_ms) *device_ms = e->eng.last_device_ms;
        if (wall_ms) *wall_ms = e->eng.last_wall_ms;
    });
}

int sdb_engine_last_launches(sdb_engine *e, int64_t *launches)
{
    return guarded([&] {
        require(e && launches, "null argument");
        *launches = e->eng.launches;
    });
}

int sdb_engine_reset(sdb_engine *e)
{
    return guarded([&] {
        require(e, "null engine");
        SDB_HIP(hipSetDevice(e->eng.device));
        e->eng.mm->reset(e->eng.stream);
        SDB_HIP(hipStreamSynchronize(e->eng.stream));
    });
}

int sdb_engine_set_fusion(sdb_engine *e, int on)
{
    return guarded([&] {
        require(e, "null engine");
        e->eng.fusion = on != 0;
    });
}

int sdb_engine_set_codegen(sdb_engine *e, int mode, int unroll)
{
    return guarded([&] {
        require(mode == SDB_CODEGEN_AUTO || mode == SDB_CODEGEN_INTERPRET || mode == SDB_CODEGEN_COMPILE,
                "unknown codegen mode");
        const int u = unroll & ~(SDB_UNROLL_QUEUED | SDB_UNROLL_BLOCK512 | SDB_UNROLL_BLOCK1024);
        require(unroll == 0 || ((u == 1 || u == 2 || u == 4) &&
                                !((unroll & SDB_UNROLL_BLOCK512) && (unroll & SDB_UNROLL_BLOCK1024))),
                "unroll must be 0, or 1, 2 or 4 (optionally | SDB_UNROLL_QUEUED and | one of "
                "SDB_UNROLL_BLOCK512, SDB_UNROLL_BLOCK1024)");
        require(e, "null engine");
        if (mode == SDB_CODEGEN_COMPILE && !jit_available())
            fail(SDB_ERR_UNSUPPORTED, "runtime compilation unavailable: " + jit_unavailable_reason());
        e->eng.codegen = mode;
        e->eng.unroll = unroll;
    });
}

int sdb_codegen_available(char *why, int why_cap)
{
    const bool ok = jit_available();
    if (why && why_cap > 0) {
        const std::string r = ok ? std::string() : jit_unavailable_reason();
        snprintf(why, (size_t)why_cap, "%s", r.c_str());
    }
    return ok ? 1 : 0;
}

int sdb_codegen_kernels_compiled(void) { return jit_kernels_compiled(); }

int sdb_codegen_tuning_report(char *buf, int cap)
{
    const std::string r = jit_tuning_report();
    if (buf && cap > 0) snprintf(buf, (size_t)cap, "%s", r.c_str());
    return (int)r.size();
}

int sdb_codegen_wait(double timeout_s) { return jit_wait_idle(timeout_s) ? 1 : 0; }

int sdb_codegen_shutdown(void)
{
    return guarded([&] { jit_shutdown(); });
}

int sdb_set_batch_graphs(int on)
{
    return guarded([&] {
        require(on >= -1 && on <= 1, "sdb_set_batch_graphs: -1, 0 or 1");
        set_batch_graphs(on);
    });
}

int sdb_batch_graph_stats(int64_t *captures, int64_t *replays)
{
    return guarded([&] { batch_graph_stats(captures, replays); });
}

int sdb_codegen_pending(void) { return jit_pending(); }

int sdb_codegen_check(const char *arch, int *count)
{
    return guarded([&] {
        require(arch && count, "null argument");
        *count = jit_check(arch);
    });
}

int sdb_engine_set_profiling(sdb_engine *e, int on)
{
    return guarded([&] {
        require(e, "null engine");
        e->eng.profile = on != 0;
    });
}

int sdb_engine_set_profile_table(sdb_engine *e, const char *table)
{
    return guarded([&] {
        require(e, "null engine");
        e->eng.prof_table = table ? table : "";
    });
}

int sdb_engine_profile_reset(sdb_engine *e)
{
    return guarded([&] {
        require(e, "null engine");
        e->eng.prof_stats.clear();
    });
}

int sdb_engine_profile_count(sdb_engine *e)
{
    return e ? (int)e->eng.prof_stats.size() : 0;
}

int sdb_engine_profile_get(sdb_engine *e, int idx, char *name, int name_cap, int64_t *launches, int64_t *rows,
                           double *ms)
{
    return guarded([&] {
        require(e && name && name_cap > 0 && launches && rows && ms, "null argument");
        require(idx >= 0 && idx < (int)e->eng.prof_stats.size(), "profile index out of range");
        auto it = e->eng.prof_stats.begin();
        std::advance(it, idx);
        std::snprintf(name, (size_t)name_cap, "%s", it->first.c_str());
        *launches = it->second.launches;
        *rows = it->second.rows;
        *ms = it->second.ms;
    });
}

uint64_t sdb_result_nrows(const sdb_result *r) { return r ? r->t.nrows : 0; }
int sdb_result_ncols(const sdb_result *r) { return r ? (int)r->t.cols.size() : 0; }

int sdb_result_col_is_u64(const sdb_result *r, int c)
{
    if (!r || c < 0 || c >= (int)r->t.cols.size()) return -1;
    return r->t.cols[c].is_u64 ? 1 : 0;
}

int sdb_result_col_copy(const sdb_result *r, int c, void *host)
{
    return guarded([&] {
        require(r && host, "null argument");
        require(c >= 0 && c < (int)r->t.cols.size(), "column out of range");
        r->t.col_to(c, host);
    });
}

int sdb_result_flags_copy(const sdb_result *r, uint8_t *host)
{
    return guarded([&] {
        require(r && host, "null argument");
        r->t.flags_to(host);
    });
}

uint64_t sdb_result_selected(const sdb_result *r) { return r ? r->t.selected() : 0; }

int sdb_result_selected_rows(const sdb_result *r, uint64_t *host)
{
    return guarded([&] {
        require(r && host, "null argument");
        r->t.selected_rows_to(host);
    });
}

int sdb_result_col_copy_selected(const sdb_result *r, int c, void *host)
{
    return guarded([&] {
        require(r && host, "null argument");
        require(c >= 0 && c < (int)r->t.cols.size(), "column out of range");
        r->t.col_selected_to(c, host);
    });
}

int sdb_result_write(const sdb_result *r, const char *path)
{
    return guarded([&] {
        require(r && path, "null argument");
        std::ofstream out(path);
        if (!out) fail(SDB_ERR_INVALID_ARG, std::string("cannot open ") + path);
        const ResultTable &t = r->t;
        const uint64_t k = t.selected();
        std::vector<uint64_t> rows(k);
        t.selected_rows_to(rows.data());
        for (uint64_t j = 0; j < k; j++) {
            // dense: the value at the row; sparse: the j-th stored value
            const uint64_t at = t.sparse ? j : rows[j];
            bool first = true;
            for (const auto &c : t.cols) {
                if (!first) out << ' ';
                if (c.is_u64) out << c.u64[at];
                else out << c.i32[at];
                first = false;
            }
            out << '\n';
        }
    });
}

void sdb_result_destroy(sdb_result *r) { delete r; }

int sdb_plan_dag_order(const char *plan_json, int32_t *order, int cap, int *count)
{
    return guarded([&] {
        require(plan_json && count, "null argument");
        const PlanResult plan = parse_plan(plan_json);
        const std::vector<int> o = dag_topological_sort(plan);
        *count = (int)o.size();
        for (int i = 0; i < (int)o.size() && i < cap; i++) order[i] = o[i];
    });
}

} // extern "C"
