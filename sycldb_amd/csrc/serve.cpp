// Serving loop over HBM-resident tables (no counterpart in the reference,
// whose application runs one query at a time: app/application.cpp).
//
// A server holds several engines (each with its own resident tables, stream
// and allocator).  Requests -- a plan for one engine -- are queued per engine
// with their arrival time.  In batching mode each engine has a worker (or,
// with single_worker, one worker takes the engine whose first waiting request
// is the oldest); when free, it admits its engine's waiting requests in
// arrival order while a slot (max_batch) and the memory budget (the sum of
// the requests' estimated footprints) allow -- always at least one -- and
// runs them as one batch (execute_batch: one shared fact scan).  Requests
// that arrive while a batch runs wait for the next one: continuous batching.
// The engines' workers overlap on the device: one engine's host work and
// kernels run beside another's batched scan, which therefore takes at most
// SERVE_BATCH_WAVES_PER_CU of each CU's wave slots; and no worker runs
// autotuner trials on requests.  A shared scan has a fixed cost that a few
// small requests do not repay, so per engine the worker learns the batch size
// from which batching pays: it keeps a running mean of a request's plain run
// time p and models a batch of n as A + B n with B a fixed fraction of p,
// learning A from every batch.  It batches an engine's n waiting requests
// only when that lowers the summed latency of all waiting requests (the
// batch's n and the m waiting for other engines); otherwise it takes only the
// oldest and runs it alone, the others staying queued in arrival order.
// Light load thus runs requests one by one, a growing queue shares scans.
// Every request's result is handed to its waiter as soon as it is ready.  In
// serial mode one worker serves all requests one at a time in arrival order
// through the plain path: the sequential baseline.
#include <algorithm>
#include <chrono>
#include <cmath>
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

// Batch cost model of the serving loop: the marginal cost of one more
// program in a shared scan, as a fraction of its plain run (measured at 600M
// rows: 0.4-0.9 ms per extra program against 1.6-2.7 ms alone), and the
// fixed part assumed before any batch has run on an engine, in plain runs
// (measured at 600M rows: ~6 ms against 1.6 ms alone).
constexpr double BATCH_MARGINAL = 0.3;
constexpr double BATCH_FIXED_INITIAL = 4.0;
// Wave slots per CU (of 32) a batched scan may take while other engines'
// workers run: a batched kernel that fills every slot makes the others'
// kernels wait for its whole scan.
constexpr int SERVE_BATCH_WAVES_PER_CU = 24;

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
    // run time and the learned fixed part of a batch's time (ms; 0 = none
    // seen yet)
    std::vector<double> plain_ms;
    std::vector<double> batch_fixed_ms;
    std::vector<std::thread> workers;

    // the queue whose first request arrived first (-1: all empty)
    int oldest_queue() const
    {
        int q = -1;
        for (int i = 0; i < (int)queues.size(); i++)
            if (!queues[i].empty() && (q < 0 || queues[i].front()->arrival < queues[q].front()->arrival)) q = i;
        return q;
    }

    // The worker of engine `own`, or (own < 0) the one worker of every
    // engine, which takes the engine whose first request is the oldest.
    void worker(int own)
    {
        // no autotuner trials on requests: their batches take the choices
        // already tuned for their kind, or the default shape
        set_batch_trials(false);
        // with a worker per engine, a batched scan leaves a quarter of each
        // CU's wave slots to the other engines' kernels
        if (own >= 0 && engines.size() > 1) set_batch_waves_per_cu(SERVE_BATCH_WAVES_PER_CU);
        while (true) {
            std::vector<std::shared_ptr<Request>> admitted;
            {
                std::unique_lock<std::mutex> lk(mu);
                auto pick = [&] { return own >= 0 ? (queues[own].empty() ? -1 : own) : oldest_queue(); };
                work_cv.wait(lk, [&] { return stopping || pick() >= 0; });
                const int q = pick();
                if (q < 0) return; // stopping with nothing left
                // a batch only when it lowers the summed latency of the
                // waiting requests; otherwise the oldest request alone, the
                // rest staying queued in arrival order with the other
                // engines' requests
                const int want = batch_pays(q, own < 0) ? max_batch : 1;
                uint64_t used = 0;
                while (!queues[q].empty() && (int)admitted.size() < want) {
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
            }
        }
    }

    // Whether engine q's waiting requests should share one scan.  With p the
    // plain run time, a batch of n costing A + B n and m requests waiting
    // for the other engines: the batch finishes all n (and lets the m start)
    // at A + B n; one by one the i-th finishes at i p and the m start after
    // n p.  Sharing pays when (A + B n)(n + m) < p n (n + 1) / 2 + n p m.
    bool batch_pays(int q, bool shared_worker) const
    {
        const double p = plain_ms[q];
        if (p <= 0) return false; // the engine's first request runs alone
        const double n = (double)std::min<size_t>(queues[q].size(), (size_t)max_batch);
        if (n < 2) return false;
        // the other engines' waiting requests: with one worker for all they
        // wait for this batch outright; with a worker per engine they still
        // share the device with it, whose launches fill every CU
        (void)shared_worker;
        double m = 0;
        for (int i = 0; i < (int)queues.size(); i++)
            if (i != q) m += (double)queues[i].size();
        const double A = batch_fixed_ms[q] > 0 ? batch_fixed_ms[q] : BATCH_FIXED_INITIAL * p;
        const double B = BATCH_MARGINAL * p;
        return (A + B * n) * (n + m) < p * n * (n + 1) / 2 + n * p * m;
    }

    // a request's result is ready: its waiter returns now, not when the
    // rest of its worker call is done
    void finish(Request &r)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            r.done = Clock::now();
            r.finished = true;
        }
        done_cv.notify_all();
    }

    // one request through the plain path (its failure is its own)
    void run_alone(Engine &eng, Request &r)
    {
        const auto t0 = Clock::now();
        r.start = t0;
        r.batch_size = 1;
        r.status = guarded([&] {
            SDB_HIP(hipSetDevice(eng.device));
            eng.mm->reset(eng.stream);
            r.result = execute_ddor_plan(eng, r.plan);
        });
        r.ran = true;
        if (r.status != SDB_OK) {
            r.error = sdb_last_error();
            finish(r);
            return;
        }
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        double &m = plain_ms[r.engine];
        m = m > 0 ? 0.8 * m + 0.2 * ms : ms;
        finish(r);
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
        bool ok = false;
        // (a group of several is admitted only for an engine whose plain
        // run time is known: batch_pays)
        if (rs.size() > 1) {
            try {
                const auto b0 = Clock::now();
                SDB_HIP(hipSetDevice(eng.device));
                std::vector<PlanResult> plans;
                for (auto &r : rs) plans.push_back(r->plan);
                std::vector<ResultTable> out = execute_batch(eng, plans, nullptr);
                for (size_t i = 0; i < rs.size(); i++) rs[i]->result = std::move(out[i]);
                ok = true;
                for (auto &r : rs) finish(*r);
                // the batch time's fixed part A, learned from every batch
                const double batch_ms = std::chrono::duration<double, std::milli>(Clock::now() - b0).count();
                const double A = std::max(0.0, batch_ms - BATCH_MARGINAL * plain_ms[e] * (double)plans.size());
                batch_fixed_ms[e] = batch_fixed_ms[e] > 0 ? 0.5 * batch_fixed_ms[e] + 0.5 * A : A;
            } catch (const std::exception &) {
                // a failing request must not fail its batch: run them one by one
            }
        }
        if (!ok)
            for (auto &r : rs)
                if (r->status == SDB_OK && !r->ran) run_alone(eng, *r);
        for (auto &r : rs)
            if (!r->finished) finish(*r);
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
        s.batch_fixed_ms.assign(n_engines, 0.0);
        // one worker per engine (each engine has its own stream, so one
        // engine's host work and kernels overlap another's), except in the
        // serial baseline and with single_worker
        if (s.serial || (opts && opts->single_worker))
            s.workers.emplace_back([&s] { s.worker(-1); });
        else
            for (int i = 0; i < n_engines; i++) s.workers.emplace_back([&s, i] { s.worker(i); });
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
