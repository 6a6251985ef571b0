// Serving loop over HBM-resident tables (no counterpart in the reference,
// whose application runs one query at a time: app/application.cpp).
//
// A server holds several engines (each with its own resident tables, stream
// and allocator).  Requests -- a plan for one engine -- are queued per engine
// with their arrival time.  In batching mode one worker, as soon as it is
// free, takes the engine whose first waiting request is the oldest and admits
// that engine's waiting requests in arrival order while a slot (max_batch)
// and the memory budget (the sum of the requests' estimated footprints) allow
// -- always at least one -- and runs them as one batch (execute_batch: one
// shared fact scan).  Requests that arrive while a batch runs wait for the
// next one: continuous batching.  One batch runs at a time: a batched launch
// fills every CU, so batches of different engines would only queue behind
// each other on the device (and a small table's requests behind a large
// table's scan).  A shared scan has a fixed cost that a few small requests do
// not repay, so per engine the worker learns the batch size from which
// batching pays: it keeps a running mean of a request's plain run time and,
// after every batch, moves the engine's threshold up when the batch took
// longer than its requests would have taken one by one and down when it took
// clearly less; admitted requests below the threshold run one by one.  In
// serial mode the same worker serves all requests one at a time in arrival
// order through the plain path: the sequential baseline.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "capi_handles.hpp"
#include "capi_util.hpp"
#include "engine.hpp"
#include "plan.hpp"

namespace sdb {

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket;
    int engine;
    PlanResult plan;
    uint64_t footprint;
    Clock::time_point arrival, start, done;
    int batch_size = 0;
    int status = SDB_OK;
    std::string error;
    ResultTable result;
    bool ran = false;      // ran through the plain path
    bool finished = false;
};

// Device memory a request's dimension side will take: its build tables
// (reference layouts and compact forms, ~12 B per key slot) and its group-by
// arrays (sums, seen, keys).  Computed from the resident tables' statistics.
uint64_t estimate_footprint(Engine &eng, const PlanResult &plan)
{
    uint64_t bytes = 1u << 20;
    const ExecutionInfo info = parse_execution_info(plan);
    for (const RelNode &rel : plan.rels) {
        if (rel.relOp != RelNodeType::TABLE_SCAN || rel.tables.size() < 2 || rel.tables[1] == "lineorder") continue;
        Table *t = eng.find_table(rel.tables[1]);
        if (!t || t->columns.empty() || !t->columns[0].loaded) continue;
        const auto mm = t->columns[0].get_min_max();
        bytes += 12ull * (uint64_t)((int64_t)mm.second - mm.first + 1);
    }
    uint64_t prod = 1;
    int keys = 0;
    for (const auto &g : info.group_by_columns) {
        Table *t = eng.find_table(g.first);
        if (!t || g.second < 0 || g.second >= (int)t->columns.size() || !t->columns[g.second].loaded) continue;
        const auto mm = t->columns[g.second].get_min_max();
        prod *= (uint64_t)((int64_t)mm.second - mm.first + 1);
        keys++;
    }
    if (keys) bytes += prod * (16ull + 4ull * keys);
    return bytes;
}

} // namespace

struct Server {
    std::vector<Engine *> engines;
    int max_batch;
    uint64_t budget;
    bool serial;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::vector<std::deque<std::shared_ptr<Request>>> queues; // per engine
    std::map<int64_t, std::shared_ptr<Request>> requests;
    int64_t next_ticket = 1;
    bool stopping = false;
    int64_t batches = 0, batched_requests = 0;
    // per engine (worker thread only): running mean of one request's plain
    // run time (ms; 0 = none seen yet) and the smallest batch worth sharing
    std::vector<double> plain_ms;
    std::vector<int> min_batch;
    std::vector<std::thread> workers;

    // the queue whose first request arrived first (-1: all empty)
    int oldest_queue() const
    {
        int q = -1;
        for (int i = 0; i < (int)queues.size(); i++)
            if (!queues[i].empty() && (q < 0 || queues[i].front()->arrival < queues[q].front()->arrival)) q = i;
        return q;
    }

    void worker()
    {
        while (true) {
            std::vector<std::shared_ptr<Request>> admitted;
            {
                std::unique_lock<std::mutex> lk(mu);
                work_cv.wait(lk, [&] { return stopping || oldest_queue() >= 0; });
                const int q = oldest_queue();
                if (q < 0) return; // stopping with nothing left
                uint64_t used = 0;
                while (!queues[q].empty() && (int)admitted.size() < max_batch) {
                    auto &r = queues[q].front();
                    if (!admitted.empty() && used + r->footprint > budget) break;
                    used += r->footprint;
                    admitted.push_back(r);
                    queues[q].pop_front();
                }
            }
            run(admitted);
            {
                std::lock_guard<std::mutex> lk(mu);
                batches++;
                batched_requests += (int64_t)admitted.size();
                for (auto &r : admitted) r->finished = true;
            }
            done_cv.notify_all();
        }
    }

    // one request through the plain path (its failure is its own)
    void run_alone(Engine &eng, Request &r)
    {
        const auto t0 = Clock::now();
        r.status = guarded([&] {
            SDB_HIP(hipSetDevice(eng.device));
            eng.mm->reset(eng.stream);
            r.result = execute_ddor_plan(eng, r.plan);
        });
        r.ran = true;
        if (r.status != SDB_OK) {
            r.error = sdb_last_error();
            return;
        }
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        double &m = plain_ms[r.engine];
        m = m > 0 ? 0.8 * m + 0.2 * ms : ms;
    }

    void run(std::vector<std::shared_ptr<Request>> &rs)
    {
        const auto start = Clock::now();
        for (auto &r : rs) {
            r->start = start;
            r->batch_size = (int)rs.size();
        }
        // all requests of one worker call share an engine (serial mode: one
        // request per call)
        const int e = rs.front()->engine;
        Engine &eng = *engines[e];
        const int n = (int)rs.size();
        bool ok = false;
        // until a plain run time is known, the first request runs alone
        size_t first_shared = plain_ms[e] > 0 ? 0 : 1;
        if (n - (int)first_shared > 1 && n - (int)first_shared >= min_batch[e]) {
            try {
                for (size_t i = 0; i < first_shared; i++) run_alone(eng, *rs[i]);
                const auto b0 = Clock::now();
                SDB_HIP(hipSetDevice(eng.device));
                std::vector<PlanResult> plans;
                for (size_t i = first_shared; i < rs.size(); i++) plans.push_back(rs[i]->plan);
                std::vector<ResultTable> out = execute_batch(eng, plans, nullptr);
                for (size_t i = first_shared; i < rs.size(); i++) rs[i]->result = std::move(out[i - first_shared]);
                ok = true;
                const double batch_ms = std::chrono::duration<double, std::milli>(Clock::now() - b0).count();
                const double alone_ms = plain_ms[e] * (double)plans.size();
                if (batch_ms > alone_ms) min_batch[e] = std::min(max_batch, (int)plans.size() + 1);
                else if (batch_ms < 0.8 * alone_ms) min_batch[e] = std::max(2, min_batch[e] - 1);
            } catch (const std::exception &) {
                // a failing request must not fail its batch: run them one by one
            }
        }
        if (!ok)
            for (auto &r : rs)
                if (r->status == SDB_OK && !r->ran) run_alone(eng, *r);
        const auto done = Clock::now();
        for (auto &r : rs) r->done = done;
    }
};

} // namespace sdb

using sdb::Server;

struct sdb_server {
    Server s;
};

extern "C" {

int sdb_server_create(sdb_engine *const *engines, int n_engines, const sdb_server_options *opts, sdb_server **out)
{
    using namespace sdb;
    return guarded([&] {
        require(engines && n_engines > 0 && out, "null argument");
        *out = nullptr;
        auto srv = std::make_unique<sdb_server>();
        Server &s = srv->s;
        for (int i = 0; i < n_engines; i++) {
            require(engines[i] != nullptr, "null engine");
            s.engines.push_back(&engines[i]->eng);
        }
        s.max_batch = opts && opts->max_batch > 0 ? opts->max_batch : 32;
        s.budget = opts && opts->memory_budget_bytes ? opts->memory_budget_bytes : (2ull << 30);
        s.serial = opts && opts->serial;
        if (s.serial) s.max_batch = 1;
        s.queues.resize(n_engines);
        s.plain_ms.assign(n_engines, 0.0);
        s.min_batch.assign(n_engines, 2);
        s.workers.emplace_back([&s] { s.worker(); });
        *out = srv.release();
    });
}

int sdb_server_submit(sdb_server *srv, int engine, const char *plan_json, int64_t *ticket)
{
    using namespace sdb;
    return guarded([&] {
        require(srv && plan_json && ticket, "null argument");
        Server &s = srv->s;
        require(engine >= 0 && engine < (int)s.engines.size(), "engine index out of range");
        auto r = std::make_shared<Request>();
        r->arrival = std::chrono::steady_clock::now();
        r->engine = engine;
        r->plan = parse_plan(plan_json);
        r->footprint = estimate_footprint(*s.engines[engine], r->plan);
        {
            std::lock_guard<std::mutex> lk(s.mu);
            require(!s.stopping, "server is stopping");
            r->ticket = s.next_ticket++;
            s.requests[r->ticket] = r;
            s.queues[engine].push_back(r);
            *ticket = r->ticket;
        }
        s.work_cv.notify_all();
    });
}

int sdb_server_wait(sdb_server *srv, int64_t ticket, sdb_result **out, sdb_request_info *info)
{
    using namespace sdb;
    std::shared_ptr<Request> r;
    const int st = guarded([&] {
        require(srv && out, "null argument");
        *out = nullptr;
        Server &s = srv->s;
        std::unique_lock<std::mutex> lk(s.mu);
        auto it = s.requests.find(ticket);
        require(it != s.requests.end(), "unknown ticket");
        r = it->second;
        s.done_cv.wait(lk, [&] { return r->finished; });
        s.requests.erase(it);
    });
    if (st != SDB_OK) return st;
    if (info) {
        using ms = std::chrono::duration<double, std::milli>;
        info->latency_ms = ms(r->done - r->arrival).count();
        info->queue_ms = ms(r->start - r->arrival).count();
        info->run_ms = ms(r->done - r->start).count();
        info->batch_size = r->batch_size;
    }
    if (r->status != SDB_OK) return guarded([&] { fail(r->status, r->error); });
    auto res = std::make_unique<sdb_result>();
    res->t = std::move(r->result);
    *out = res.release();
    return SDB_OK;
}

int sdb_server_stats(sdb_server *srv, int64_t *batches, int64_t *requests)
{
    using namespace sdb;
    return guarded([&] {
        require(srv, "null server");
        std::lock_guard<std::mutex> lk(srv->s.mu);
        if (batches) *batches = srv->s.batches;
        if (requests) *requests = srv->s.batched_requests;
    });
}

void sdb_server_destroy(sdb_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->s.mu);
        srv->s.stopping = true;
    }
    srv->s.work_cv.notify_all();
    for (auto &t : srv->s.workers) t.join();
    delete srv;
}

} // extern "C"
