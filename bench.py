"""Benchmark of the SYCLDB operator hot path on one MI355X.

Workload (default "c3"): the q2.x star join on a 600,000,000-row lineorder
table resident in HBM -- part filtered to one category (4 %), supplier to one
region (20 %), joined with date, GROUP BY (year, brand) with an exact int64
SUM(revenue) -- run as the reference's executor runs it (TransientTable
bundles flushed at joins and at the aggregate), with each bundle fused into
one kernel.  One step = one full plan execution (dimension filters, hash-table
builds, the fact-table pass, the group-by and the result download).

Prints one JSON line: fact rows/s over the timed steps, the fused fact
kernel's achieved HBM bandwidth against the 8 TB/s peak (HIP events around
every launch of that kernel), the unfused per-operator path for comparison,
and the CPU oracle timed on a bounded sample of the same table on this host.
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_CAL_BYTES = 1 << 30
# HBM traffic of the fused fact kernel per launch, from the committed
# rocprofv3 counter summary of this workload (FETCH_SIZE + WRITE_SIZE scaled by
# the copy-kernel calibration); reported as roofline.traffic
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "c3_q2x_600M_fused_fact_kernel.txt")

WORKLOADS = {
    # name: (fact rows, plan builder, fact columns the plan reads)
    "c3": (600_000_000, "config3_plan", (3, 4, 5, 12)),
    "c2": (60_000_000, "config2_plan", (2, 4, 5, 12)),
    "c1": (6_000_000, "config1_plan", (5, 8, 9, 11)),
}
DESCRIPTIONS = {
    "c3": "600M-row lineorder resident in HBM; part category=1 (4%), supplier region=1 (20%), date; "
          "GROUP BY d_year, p_brand1; SUM(lo_revenue) (q2.x)",
    "c2": "60M-row lineorder; customer and supplier region=2, years 1992-1997; GROUP BY c_nation, s_nation, "
          "d_year; SUM(lo_revenue) (q3.1)",
    "c1": "6M-row lineorder; semi-join date year=1993, discount 1..3, quantity<25; SUM(extendedprice*discount) "
          "(q1.1)",
}


class no_gc:
    """Timed regions run with Python's cyclic garbage collector paused (after
    a collection): a collection pass over the interpreter's objects takes
    milliseconds and would land at random in a timed step."""

    def __enter__(self):
        gc.collect()
        gc.disable()

    def __exit__(self, *exc):
        gc.enable()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batched_measure(eng, n_pipes, seed, reps):
    """32 (by default) pipelines from the four templates with seeded random
    constants over the resident tables: run one by one through the plain
    path, then as one batch sharing the fact scans.  Results must agree."""
    import torch
    from sycldb_amd import plans as PL
    pipes = PL.random_pipelines(n_pipes, seed)
    plans = [PL.to_json(PL.build(t, **p)) for t, p in pipes]  # JSON text, as the C ABI takes plans
    alone = [eng.execute(p).rows() for p in plans]  # the reference results
    # warm: every plan shape's kernels compile and the autotuner settles
    # before anything is timed
    from sycldb_amd.engine import codegen_pending, codegen_tuning_report, codegen_wait
    for _ in range(64):
        for p in plans:
            eng.execute(p)
        if codegen_pending() == 0:
            break
        codegen_wait(300)
    with no_gc():
        for p in plans:
            eng.execute(p)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for p in plans:
                eng.execute(p)
        torch.cuda.synchronize()
        seq_s = (time.perf_counter() - t0) / reps
    # warm: the batched kernel's autotuner tries its launch shapes (and
    # compiles them) on the first batches
    for k in range(16):
        results, nb = eng.execute_batch(plans)
        assert [r.rows() for r in results] == alone, "batched results differ from running each plan alone"
        if k >= 1 and codegen_pending() == 0:
            break
    torch.cuda.synchronize()
    from sycldb_amd.engine import batch_graph_stats, set_batch_graphs

    def time_batches(warm):
        with no_gc():
            for _ in range(warm):
                eng.execute_batch(plans)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                eng.execute_batch(plans)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps
    # the launches as plain stream operations (the default), then as
    # replayed hipGraphs; each mode warmed first, right before its timing
    set_batch_graphs(False)
    bat_s = time_batches(1)
    set_batch_graphs(True)
    eng.execute_batch(plans)
    c0, r0 = batch_graph_stats()
    graph_s = time_batches(1)
    c1, r1 = batch_graph_stats()
    set_batch_graphs(None)
    # per template: its 8 pipelines one by one and as one batch
    per_template = {}
    for tmpl in ("t1", "t2", "t3", "t4"):
        sub = [plans[i] for i, (t, _) in enumerate(pipes) if t == tmpl]
        if len(sub) < 2:
            continue
        sub_alone = [alone[i] for i, (t, _) in enumerate(pipes) if t == tmpl]
        # warm: the autotuner tunes each kind of batch (its mask tables)
        # on its first launches
        for k in range(16):
            results, _ = eng.execute_batch(sub)
            assert [r.rows() for r in results] == sub_alone, "batched %s results differ" % tmpl
            if k >= 1 and codegen_pending() == 0:
                break
        torch.cuda.synchronize()
        with no_gc():
            for p in sub:
                eng.execute(p)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                for p in sub:
                    eng.execute(p)
            torch.cuda.synchronize()
            t_seq = (time.perf_counter() - t0) / reps
            eng.execute_batch(sub)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                eng.execute_batch(sub)
            torch.cuda.synchronize()
            t_bat = (time.perf_counter() - t0) / reps
        per_template[tmpl] = {"pipelines": len(sub), "sequential_ms": t_seq * 1e3, "batched_ms": t_bat * 1e3,
                              "speedup": t_seq / t_bat}
    log("batched kernel tuning:\n" + "\n".join(l for l in codegen_tuning_report().splitlines() if "batched" in l))
    return {"pipelines": n_pipes, "pipeline_seed": seed, "sharing_the_scan": nb,
            "sequential_ms": seq_s * 1e3, "batched_ms": bat_s * 1e3,
            "sequential_pipelines_per_s": n_pipes / seq_s, "batched_pipelines_per_s": n_pipes / bat_s,
            "speedup": seq_s / bat_s, "results_identical": True,
            "launch": "stream operations (copy, mask build, batched kernel)",
            "batched_ms_graph_replay": graph_s * 1e3, "per_template": per_template,
            "graph": "%d captures, %d replays while timed" % (c1 - c0, r1 - r0)}


def served_measure(eng_big, seed, n_requests, rates, max_batch):
    """Serving: three resident fact tables (6M, 60M and the benchmark's
    600M rows) with their dimensions; Poisson arrivals at a fixed rate of
    pipelines from the four templates with random constants and table sizes.
    At every rate the same trace is served one request at a time in arrival
    order (the plain sequential baseline), with continuous batching by one
    worker per engine (the engines overlap on the device), and with
    continuous batching by one worker for all engines; every result of the
    batched runs must equal the sequential one."""
    import numpy as np
    from sycldb_amd import plans as PL
    from sycldb_amd.engine import Engine, Server, replay
    # a fixed launch configuration: no autotuning trials while serving
    engines = []
    for i, n in enumerate((6_000_000, 60_000_000)):
        e = Engine(unroll=2)
        e.generate(n, seed + 10 + i)
        engines.append(e)
    eng_big.set_codegen("auto", 2)
    engines.append(eng_big)
    try:
        rng = np.random.default_rng(seed)
        pipes = PL.random_pipelines(n_requests, seed + 2)
        reqs = [(int(rng.integers(0, 3)), PL.build(t, **p)) for t, p in pipes]
        gaps = rng.exponential(1.0, n_requests)
        # compile every kernel the trace needs before measuring: each plan
        # alone, and each engine's requests as one batch
        for e in range(3):
            sub = [plan for ei, plan in reqs if ei == e]
            for plan in sub:
                engines[e].execute(plan)
            for k in range(0, len(sub), max_batch):
                if len(sub[k:k + max_batch]) > 1:
                    engines[e].execute_batch(sub[k:k + max_batch])

        def serve(serial, arrivals, single_worker=False):
            with Server(engines, max_batch=max_batch, serial=serial, single_worker=single_worker) as srv, no_gc():
                t0 = time.perf_counter()
                out = replay(srv, reqs, arrivals)
                wall = time.perf_counter() - t0
                stats = srv.stats()
            lat = np.array([info["latency_ms"] for _, info in out])
            return out, {"pipelines_per_s": n_requests / wall, "p50_ms": float(np.percentile(lat, 50)),
                         "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
                         "batches": stats["batches"], "mean_batch": stats["requests"] / max(stats["batches"], 1),
                         "wall_s": wall}

        by_rate = {}
        for qps in rates:
            arrivals = np.cumsum(gaps / qps)
            out_seq, seq = serve(True, arrivals)
            out_one, one = serve(False, arrivals, single_worker=True)
            out_bat, bat = serve(False, arrivals)
            identical = all(a.rows() == b.rows() == c.rows()
                            for (a, _), (b, _), (c, _) in zip(out_seq, out_bat, out_one))
            assert identical, "served results differ from the sequential ones"
            for m in (seq, one, bat):
                m["of_offered"] = m["pipelines_per_s"] / qps
            by_rate["%g" % qps] = {"sequential": seq, "batched": bat, "batched_one_worker": one,
                                   "throughput_ratio": bat["pipelines_per_s"] / seq["pipelines_per_s"],
                                   "p50_ratio": seq["p50_ms"] / bat["p50_ms"], "p99_ratio": seq["p99_ms"] / bat["p99_ms"],
                                   "results_identical": identical}
        return {"requests": n_requests, "tables_rows": [6_000_000, 60_000_000, eng_big_rows(eng_big)],
                "max_batch": max_batch, "by_qps": by_rate}
    finally:
        for e in engines[:2]:
            e.close()


def config_measure(name, seed, steps, cpu_seconds):
    """One of the smaller plain configurations at its own size: q1.1 on the
    6M-row table (C1) or the q3.1 join on the 60M-row table (C2).  GPU: the
    fused plan, settled first, then timed steps, then as many steps with HIP
    events around the fact kernel for its launch time.  CPU: the oracle (the
    reference's unfused operator sequence) on the same seeded table at full
    size, on this host's cores.  The two results must agree."""
    import torch
    from oracle import datagen, oracle as O, plan_oracle as PO
    from sycldb_amd import plans as PL, schema as S
    from sycldb_amd.engine import Engine, codegen_pending, codegen_wait
    rows, builder, cols = WORKLOADS[name]
    plan = getattr(PL, builder)()
    text = PL.to_json(plan)
    out = {"workload": DESCRIPTIONS[name], "fact_rows": rows, "algorithmic_bytes_per_row": 4 * len(cols)}
    with Engine() as eng:
        eng.generate(rows, seed, fact_columns=cols)
        res = eng.execute(text)
        runs = 1
        while codegen_pending() > 0 and runs < 64:
            codegen_wait(300)
            res = eng.execute(text)
            runs += 1
        with no_gc():
            for _ in range(10):
                eng.execute(text)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                eng.execute(text)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        # the fact kernel's launch time from HIP events around its launches,
        # in steps of their own: at these sizes the events' host time is a
        # visible part of a step
        eng.set_profiling(True, table="lineorder")
        eng.profile_reset()
        for _ in range(steps):
            eng.execute(text)
        torch.cuda.synchronize()
        launches, krows, kms = eng.profile().get("fused:lineorder", (0, 0, 0.0))
        eng.set_profiling(False)
        gpu_rows = res.rows()
    out["gpu"] = {"steps": steps, "ms_per_step": wall * 1e3 / steps, "fact_rows_per_s": rows * steps / wall,
                  "kernel_ms": kms / max(launches, 1),
                  "kernel_hbm_frac": (4 * len(cols) * krows / (kms / 1e3)) / (HBM_PEAK_GBS * 1e9) if kms else None}
    tables = datagen.generate_tables(S, rows, seed, fact_columns=cols)
    want = PO.execute_plan(plan, tables).rows()
    assert want == gpu_rows, "%s: GPU and CPU oracle results differ" % name
    reps, elapsed = 0, 0.0
    while elapsed < cpu_seconds or reps < 2:
        t0 = time.perf_counter()
        PO.execute_plan(plan, tables)
        elapsed += time.perf_counter() - t0
        reps += 1
    out["cpu_baseline"] = {"fact_rows_per_s": rows * reps / elapsed, "ms_per_step": elapsed * 1e3 / reps,
                           "cores": O.num_threads(), "kind": "port", "reps": reps}
    out["results_identical"] = True
    out["gpu_over_cpu"] = out["gpu"]["fact_rows_per_s"] / out["cpu_baseline"]["fact_rows_per_s"]
    return out


def eng_big_rows(eng):
    return getattr(eng, "fact_rows", None)


def measured_traffic(workload, rows):
    """HBM bytes per fused fact-kernel launch from the committed counter
    summary, when it covers this workload at this size."""
    if workload != "c3" or rows != WORKLOADS["c3"][0] or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    vals = {}
    for line in open(TRAFFIC_SUMMARY):
        if ":" in line and line.split(":")[0].strip() in ("traffic_bytes_per_launch", "launch_rows"):
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v.split()[0])
    if "traffic_bytes_per_launch" not in vals:
        return None
    return {"bytes_per_launch": vals["traffic_bytes_per_launch"],
            "source": os.path.relpath(TRAFFIC_SUMMARY, ROOT)}


def cpu_baseline(plan, fact_cols, sample_rows, seed):
    """Times the CPU oracle (C/OpenMP operator restatement driven by the plan
    restatement, unfused like the reference's CPU path) on the first
    `sample_rows` rows of the same seeded table."""
    from oracle import datagen, oracle as O, plan_oracle as PO
    from sycldb_amd import schema as S
    t0 = time.time()
    tables = datagen.generate_tables(S, sample_rows, seed, fact_columns=fact_cols)
    log("cpu sample generated in %.1fs" % (time.time() - t0))
    PO.execute_plan(plan, tables)  # warm
    reps, elapsed = 0, 0.0
    while elapsed < 8.0 or reps < 2:
        t0 = time.perf_counter()
        PO.execute_plan(plan, tables)
        elapsed += time.perf_counter() - t0
        reps += 1
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": sample_rows * reps / elapsed, "unit": "fact rows/s", "cores": O.num_threads(),
            "kind": "port",
            "sample": "first %d rows of the same seeded lineorder (full dimensions), %d timed repetitions, "
                      "%.1f s of CPU work" % (sample_rows, reps, elapsed),
            "nproc": os.cpu_count(), "cpu_model": cpu_model,
            "host_copy_GBps": round(O.host_copy_gbps(1 << 30), 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=11)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--fact-rows", type=int, default=0)
    ap.add_argument("--seed", type=int, default=45)
    ap.add_argument("--cpu-sample-rows", type=int, default=60_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-unfused", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-pipelines", type=int, default=32)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the q1.1 (6M rows) and q3.1 (60M rows) configurations at their own sizes")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-qps", default="1000,3000",
                    help="arrival rates (requests/s), comma separated")
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--codegen", default="auto", choices=["auto", "interpret", "compile"])
    ap.add_argument("--unroll", type=int, default=0, choices=[0, 1, 2, 4])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB three times (reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the operator path has no multi-GPU form (the fact table would be "
              "sharded by row ranges with replicated dimension tables, and the int64 partial results "
              "reduced across GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        x = torch.ones(COPY_CAL_BYTES // 4, dtype=torch.int32, device="cuda")
        for _ in range(3):
            y = x.clone()
        torch.cuda.synchronize()
        log("copied %d bytes three times" % (y.numel() * 4))
        return
    from sycldb_amd import plans as PL
    from sycldb_amd.engine import (Engine, codegen_kernels_compiled, codegen_pending, codegen_tuning_report,
                                   codegen_wait)

    rows, builder, fact_cols = WORKLOADS[args.workload]
    rows = args.fact_rows or rows
    plan = getattr(PL, builder)()
    # the plan as the C ABI takes it (JSON text), serialised once, as a
    # client that runs a query repeatedly holds it
    plan_text = PL.to_json(plan)
    eng = Engine(codegen=args.codegen, unroll=args.unroll)
    t0 = time.time()
    # all fact columns: the batched and served pipelines read up to nine
    eng.generate(rows, args.seed)
    torch.cuda.synchronize()
    log("generated %d fact rows on the device in %.1fs" % (rows, time.time() - t0))

    # The library compiles a new plan's kernels on a background thread (the
    # plan runs on the interpreter kernel meanwhile) and then tunes the fact
    # kernel's launch configuration: executions continue, untimed, until
    # nothing is pending, whatever --warmup says.  Then the warmup steps.
    t_tune = time.perf_counter()
    res = eng.execute(plan_text)
    tuning_runs = 1
    while codegen_pending() > 0 and tuning_runs < 64:
        codegen_wait(300)
        res = eng.execute(plan_text)
        tuning_runs += 1
    torch.cuda.synchronize()
    tuning_s = time.perf_counter() - t_tune
    if codegen_pending() > 0:
        log("warning: the autotuner has not settled after %d executions" % tuning_runs)
    groups = len(res.rows())
    tuning = codegen_tuning_report()
    if tuning:
        log("autotuner (%d executions, %.2f s):\n%s" % (tuning_runs, tuning_s, tuning))
    # HIP events around the fact-table launches only (the dimension
    # bundles run untimed, without the events' host cost)
    eng.set_profiling(True, table="lineorder")
    dev_ms = []
    with no_gc():
        # the warmup steps right before the timed ones, after the set-up
        # above: the device idling meanwhile (logging, a collection) made
        # the first steps after it up to 0.2 ms slower
        for _ in range(args.warmup):
            eng.execute(plan_text)
        eng.profile_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.execute(plan_text)
            dev_ms.append(eng.last_timing()[0])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    prof = eng.profile()
    eng.set_profiling(False)
    ms_per_step = wall * 1000.0 / args.steps
    value = rows * args.steps / wall
    log("fused: %.3f ms/step, %.3e rows/s, %d groups; profile %s" % (ms_per_step, value, groups, prof))

    launches, krows, kms = prof.get("fused:lineorder", (0, 0, 0.0))
    bytes_per_row = 4 * len(fact_cols)
    achieved = (bytes_per_row * krows / (kms / 1000.0)) / 1e9 if kms > 0 else None
    traffic = measured_traffic(args.workload, rows)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "kernel": "sdb_fused_program (fused lineorder bundle, compiled for its program shape)" if args.codegen != "interpret" else "k_fused (fused lineorder bundle, interpreted)", "launches_per_step": launches // max(args.steps, 1),
                "avg_launch_ms": kms / launches if launches else None,
                "algorithmic_bytes_per_row": bytes_per_row}

    unfused = None
    if not args.no_unfused:
        eng.set_fusion(False)
        eng.execute(plan_text)
        t0 = time.perf_counter()
        n_unf = max(2, args.steps // 3)
        for _ in range(n_unf):
            r2 = eng.execute(plan_text)
        torch.cuda.synchronize()
        unf_ms = (time.perf_counter() - t0) * 1000.0 / n_unf
        assert r2.rows() == res.rows(), "unfused and fused results differ"
        unfused = {"ms_per_step": unf_ms, "value": rows / (unf_ms / 1000.0), "launches": eng.last_launches()}
        eng.set_fusion(True)
        log("unfused: %.3f ms/step" % unf_ms)

    batched = None
    if not args.no_batched:
        batched = batched_measure(eng, args.batch_pipelines, args.seed + 1, args.batch_reps)
        log("batched: %s" % batched)

    served = None
    if not args.no_served:
        eng.fact_rows = rows
        served = served_measure(eng, args.seed + 3, args.serve_requests,
                                [float(q) for q in str(args.serve_qps).split(",")], args.serve_max_batch)
        log("served: %s" % served)

    configs = None
    if not args.no_configs and args.workload == "c3":
        configs = {}
        for name in ("c1", "c2"):
            configs[name] = config_measure(name, args.seed, 100, 3.0 if args.no_cpu else 6.0)
            log("%s: %s" % (name, configs[name]))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(plan, fact_cols, min(args.cpu_sample_rows, rows), args.seed)
        log("cpu baseline: %s" % cpu)

    out = {
        "metric": "fact rows/s + % HBM roofline at 600M rows; batched vs plain pipelines/s; p99",
        "value": value,
        "unit": "fact rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded star schema generated on the device (SplitMix64 counter hash)",
        "config": {"workload": DESCRIPTIONS[args.workload], "fact_rows": rows, "plan": args.workload,
                   "device_ms_per_step_median": sorted(dev_ms)[len(dev_ms) // 2], "groups": groups,
                   "fusion": True, "codegen": args.codegen, "unroll": args.unroll,
                   "kernels_compiled": codegen_kernels_compiled(), "tuning": tuning.strip().splitlines(),
                   "tuning_runs": tuning_runs, "tuning_s": tuning_s,
                   "unfused": unfused},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
