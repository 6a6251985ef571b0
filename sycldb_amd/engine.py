"""Engine API: HBM-resident star-schema tables and plan execution.

Wraps the C ABI's engine (include/sycldb.h): tables are created on the device,
filled by the seeded device generator (or uploaded / loaded from the
reference's columnar binary files), and plans -- JSON documents mirroring
calciteserver.thrift -- run through the TransientTable / KernelBundle model
(models/transient_table.hpp) to completion.
"""
import ctypes
import json
import weakref

import numpy as np

from . import _lib as L
from . import schema as S


class Result:
    """Host copy of a plan's final table (TransientTable after the last
    flush): ``nrows`` rows, selection ``flags``, non-null ``columns``.

    The table stays in the library's host memory; the selected rows (what
    result_writer.cpp:41-67 writes) and the full ``flags`` and ``columns``
    arrays are copied into numpy on first access."""

    def __init__(self, handle):
        lib = L.lib
        self._h = ctypes.c_void_p(handle.value if isinstance(handle, ctypes.c_void_p) else handle)
        self._fin = weakref.finalize(self, lib.sdb_result_destroy, self._h)
        self.nrows = int(lib.sdb_result_nrows(self._h))
        self._ncols = int(lib.sdb_result_ncols(self._h))
        self._is_u64 = None
        self._selected = None
        self._flags = None
        self._columns = None

    def _u64(self):
        if self._is_u64 is None:
            self._is_u64 = [L.lib.sdb_result_col_is_u64(self._h, c) == 1 for c in range(self._ncols)]
        return self._is_u64

    def _fetch_selected(self):
        if self._selected is None:
            lib = L.lib
            k = int(lib.sdb_result_selected(self._h))
            rows = np.zeros(k, dtype=np.uint64)
            if k:
                L.check(lib.sdb_result_selected_rows(self._h, rows.ctypes.data_as(ctypes.c_void_p)))
            cols = []
            for c, u64 in enumerate(self._u64()):
                arr = np.zeros(k, dtype=np.uint64 if u64 else np.int32)
                if k:
                    L.check(lib.sdb_result_col_copy_selected(self._h, c, arr.ctypes.data_as(ctypes.c_void_p)))
                cols.append(arr)
            self._selected = (rows, cols)
        return self._selected

    @property
    def selected_rows(self):
        """Row numbers of the selected rows, ascending."""
        return self._fetch_selected()[0]

    @property
    def selected_columns(self):
        """Each column's values at the selected rows."""
        return self._fetch_selected()[1]

    @property
    def flags(self):
        if self._flags is None:
            f = np.zeros(self.nrows, dtype=np.uint8)
            if self.nrows:
                L.check(L.lib.sdb_result_flags_copy(self._h, f.ctypes.data_as(ctypes.c_void_p)))
            self._flags = f
        return self._flags

    @property
    def columns(self):
        if self._columns is None:
            cols = []
            for c, u64 in enumerate(self._u64()):
                arr = np.zeros(self.nrows, dtype=np.uint64 if u64 else np.int32)
                if self.nrows:
                    L.check(L.lib.sdb_result_col_copy(self._h, c, arr.ctypes.data_as(ctypes.c_void_p)))
                cols.append((u64, arr))
            self._columns = cols
        return self._columns

    def rows(self):
        """Selected rows as tuples of Python ints (result_writer.cpp:41-67)."""
        cols = [c.tolist() for c in self.selected_columns]
        return list(zip(*cols)) if cols else [() for _ in range(len(self.selected_rows))]

    def text(self):
        return "".join(" ".join(str(v) for v in row) + "\n" for row in self.rows())


CODEGEN_MODES = {"auto": L.CODEGEN_AUTO, "interpret": L.CODEGEN_INTERPRET, "compile": L.CODEGEN_COMPILE}
UNROLL_QUEUED = 16  # SDB_UNROLL_QUEUED


def codegen_available():
    """(available, reason): whether fused kernels can be compiled at run time."""
    why = ctypes.create_string_buffer(512)
    ok = L.lib.sdb_codegen_available(why, 512) == 1
    return ok, why.value.decode()


def codegen_kernels_compiled():
    return int(L.lib.sdb_codegen_kernels_compiled())


def codegen_wait(timeout_s=600.0):
    """Waits until every fused kernel queued for the background compiler so
    far is compiled; False if timeout_s passed first."""
    return L.lib.sdb_codegen_wait(float(timeout_s)) == 1


def codegen_pending():
    """Program shapes still compiling or still tuning (plus batched-kernel
    libraries still tuning); 0 once every launch seen runs its final
    configuration."""
    return int(L.lib.sdb_codegen_pending())


def set_batch_graphs(on):
    """Batched launches as replayed hipGraphs (True), as plain stream
    operations (False), or the default (None: off unless SDB_BATCH_GRAPH=1)."""
    L.check(L.lib.sdb_set_batch_graphs(-1 if on is None else int(bool(on))))


def batch_graph_stats():
    """(captures, replays) of batched-launch graphs so far."""
    c, r = ctypes.c_int64(), ctypes.c_int64()
    L.check(L.lib.sdb_batch_graph_stats(ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def codegen_tuning_report():
    """The autotuner's choice per program shape (text, one line each)."""
    n = L.lib.sdb_codegen_tuning_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.lib.sdb_codegen_tuning_report(buf, n + 1)
    return buf.value.decode()


class Engine:
    """One device's tables and executor.

    fusion:  each KernelBundle runs as one fused kernel (True) or one kernel
             per operator (False, the reference's own mode).
    codegen: how a fused kernel is produced -- "auto" (compiled at run time
             for the program's shape when hipRTC loads, else interpreted),
             "interpret" or "compile".
    unroll:  4-row chunks per thread per iteration of a compiled kernel
             (1, 2, 4), or 0: chosen per program shape by the autotuner.
             1, 2 or 4 | UNROLL_QUEUED also runs a program's row tail from
             a per-wave queue drained 64 rows at a time.
    """

    def __init__(self, device_budget_bytes=0, segment_size=0, fusion=True, device=0, codegen="auto", unroll=0):
        opts = L.EngineOptions(device_budget_bytes, segment_size, 0 if fusion else 1, device)
        h = ctypes.c_void_p()
        L.check(L.lib.sdb_engine_create(ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self.segment_size = segment_size or (1 << 30)
        self.fusion = fusion
        try:
            self.set_codegen(codegen, unroll)
        except Exception:
            self.close()
            raise

    def set_codegen(self, mode="auto", unroll=0):
        if mode not in CODEGEN_MODES:
            raise ValueError("codegen must be one of %s" % sorted(CODEGEN_MODES))
        L.check(L.lib.sdb_engine_set_codegen(self._h, CODEGEN_MODES[mode], int(unroll)))
        self.codegen, self.unroll = mode, unroll

    def close(self):
        if self._h:
            L.lib.sdb_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def stream(self):
        return L.lib.sdb_engine_stream(self._h)

    # ---------------------------------------------------------------- tables
    def create_table(self, name, nrows):
        L.check(L.lib.sdb_engine_table_create(self._h, name.encode(), int(nrows)))

    def column_ptr(self, name, col):
        p = ctypes.c_void_p()
        L.check(L.lib.sdb_engine_column_alloc(self._h, name.encode(), int(col), ctypes.byref(p)))
        return p.value

    def upload(self, name, col, data):
        a = np.ascontiguousarray(data, dtype=np.int32)
        L.check(L.lib.sdb_engine_column_upload(self._h, name.encode(), int(col), a.ctypes.data_as(ctypes.c_void_p)))

    def download(self, name, col, nrows):
        a = np.empty(int(nrows), dtype=np.int32)
        L.check(L.lib.sdb_engine_column_download(self._h, name.encode(), int(col),
                                                 a.ctypes.data_as(ctypes.c_void_p)))
        return a

    def finalize(self, name):
        L.check(L.lib.sdb_engine_table_finalize(self._h, name.encode()))

    def load_dir(self, name, directory):
        L.check(L.lib.sdb_engine_table_load_dir(self._h, name.encode(), directory.encode()))

    def load_tables(self, tables):
        """tables: {name: {col: int32 array}} (host data)."""
        for name, cols in tables.items():
            n = len(next(iter(cols.values())))
            self.create_table(name, n)
            for c, data in cols.items():
                self.upload(name, c, data)
            self.finalize(name)

    def generate(self, fact_rows, seed, sizes=None, fact_columns=None, tables=None):
        """Generates the seeded star schema on the device (schema.column_spec)."""
        sizes = dict(S.DIM_SIZES, **(sizes or {}))
        s = self.stream
        tables = tables or ("ddate", "customer", "supplier", "part", "lineorder")
        for t in tables:
            n = fact_rows if t == "lineorder" else sizes[t]
            self.create_table(t, n)
            cols = S.generation_order(t)
            if t == "lineorder" and fact_columns is not None:
                cols = [c for c in cols if c in fact_columns]
            for c in cols:
                kind = S.column_spec(t, c, sizes)
                out = ctypes.c_void_p(self.column_ptr(t, c))
                sid = S.stream_id(t, c)
                if kind[0] == "rowid":
                    L.check(L.lib.sdb_gen_rowid(out, n, 0, s))
                elif kind[0] == "uniform":
                    L.check(L.lib.sdb_gen_uniform(out, n, seed, sid, 0, kind[1], kind[2], s))
                elif kind[0] == "child":
                    base = ctypes.c_void_p(self.column_ptr(t, kind[1]))
                    L.check(L.lib.sdb_gen_child(out, base, n, seed, sid, 0, kind[2], kind[3], kind[4], s))
                elif kind[0] == "lookup":
                    table = ctypes.c_void_p(self.column_ptr(kind[1], kind[2]))
                    L.check(L.lib.sdb_gen_lookup(out, n, seed, sid, 0, table, sizes[kind[1]], s))
                elif kind[0] == "date":
                    pass  # filled below from the host calendar
                else:
                    raise ValueError(kind)
            if t == "ddate":
                dk, yr, ym = date_columns(sizes["ddate"])
                self.upload(t, S.D_DATEKEY, dk)
                self.upload(t, S.D_YEAR, yr)
                self.upload(t, S.D_YEARMONTHNUM, ym)
            self.finalize(t)

    # ----------------------------------------------------------- execution
    def execute_batch(self, plans):
        """Runs several plans as one batch: their fact-table passes share one
        scan.  Returns (results in plan order, number that shared the scan);
        each result equals the plan's result when run alone."""
        texts = [p if isinstance(p, str) else json.dumps(p) for p in plans]
        n = len(texts)
        arr = (ctypes.c_char_p * max(n, 1))(*[t.encode() for t in texts])
        out = (ctypes.c_void_p * max(n, 1))()
        batched = ctypes.c_int(0)
        L.check(L.lib.sdb_engine_execute_batch(self._h, arr, n, out, ctypes.byref(batched)))
        return [Result(out[i]) for i in range(n)], batched.value

    def execute(self, plan):
        text = plan if isinstance(plan, str) else json.dumps(plan)
        h = ctypes.c_void_p()
        L.check(L.lib.sdb_engine_execute(self._h, text.encode(), ctypes.byref(h)))
        return Result(h)

    def last_timing(self):
        d, w = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib.sdb_engine_last_timing(self._h, ctypes.byref(d), ctypes.byref(w)))
        return d.value, w.value

    def last_launches(self):
        n = ctypes.c_int64()
        L.check(L.lib.sdb_engine_last_launches(self._h, ctypes.byref(n)))
        return n.value

    def reset(self):
        L.check(L.lib.sdb_engine_reset(self._h))

    def set_fusion(self, on=True):
        L.check(L.lib.sdb_engine_set_fusion(self._h, 1 if on else 0))
        self.fusion = bool(on)

    def set_profiling(self, on=True, table=None):
        """Per-launch timing of the bundles (all of them, or those of `table`)."""
        L.check(L.lib.sdb_engine_set_profile_table(self._h, table.encode() if table else None))
        L.check(L.lib.sdb_engine_set_profiling(self._h, 1 if on else 0))

    def profile_reset(self):
        L.check(L.lib.sdb_engine_profile_reset(self._h))

    def profile(self):
        """{name: (launches, rows, ms)} accumulated since profile_reset()."""
        out = {}
        buf = ctypes.create_string_buffer(256)
        for i in range(L.lib.sdb_engine_profile_count(self._h)):
            la, ro, ms = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
            L.check(L.lib.sdb_engine_profile_get(self._h, i, buf, 256, ctypes.byref(la), ctypes.byref(ro),
                                                 ctypes.byref(ms)))
            out[buf.value.decode()] = (la.value, ro.value, ms.value)
        return out


def date_columns(ndays):
    """Date dimension from 1992-01-01: datekey yyyymmdd, year, yearmonthnum."""
    import datetime
    d0 = datetime.date(1992, 1, 1)
    dk = np.empty(ndays, np.int32)
    yr = np.empty(ndays, np.int32)
    ym = np.empty(ndays, np.int32)
    for i in range(ndays):
        d = d0 + datetime.timedelta(days=i)
        dk[i] = d.year * 10000 + d.month * 100 + d.day
        yr[i] = d.year
        ym[i] = d.year * 100 + d.month
    return dk, yr, ym


def dag_order(plan):
    """dag_topological_sort of the plan (host only)."""
    text = plan if isinstance(plan, str) else json.dumps(plan)
    order = (ctypes.c_int32 * 4096)()
    count = ctypes.c_int()
    L.check(L.lib.sdb_plan_dag_order(text.encode(), order, 4096, ctypes.byref(count)))
    return list(order[:count.value])


class Server:
    """Serving loop over engines with resident tables (sdb_server_*):
    requests are plans for one engine; with batching, each engine's worker
    (one per engine, so engines overlap on the device; single_worker=True:
    one for all) admits the waiting requests (slots and memory budget
    permitting) and runs them as one batch sharing the fact scan, while later
    arrivals queue for the next batch; serial=True runs every request alone
    in arrival order.
    The engines must not be used directly while the server runs."""

    def __init__(self, engines, max_batch=32, memory_budget_bytes=0, serial=False, single_worker=False):
        self._engines = list(engines)
        arr = (ctypes.c_void_p * len(self._engines))(*[e._h for e in self._engines])
        opts = L.ServerOptions(max_batch, memory_budget_bytes, 1 if serial else 0, 1 if single_worker else 0)
        h = ctypes.c_void_p()
        L.check(L.lib.sdb_server_create(arr, len(self._engines), ctypes.byref(opts), ctypes.byref(h)))
        self._h = h

    def submit(self, engine_index, plan):
        text = plan if isinstance(plan, str) else json.dumps(plan)
        t = ctypes.c_int64()
        L.check(L.lib.sdb_server_submit(self._h, int(engine_index), text.encode(), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        """(Result, info dict with latency_ms, queue_ms, run_ms, batch_size)."""
        h = ctypes.c_void_p()
        info = L.RequestInfo()
        L.check(L.lib.sdb_server_wait(self._h, int(ticket), ctypes.byref(h), ctypes.byref(info)))
        return Result(h), {"latency_ms": info.latency_ms, "queue_ms": info.queue_ms, "run_ms": info.run_ms,
                           "batch_size": info.batch_size}

    def stats(self):
        b, r = ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib.sdb_server_stats(self._h, ctypes.byref(b), ctypes.byref(r)))
        return {"batches": b.value, "requests": r.value}

    def close(self):
        if self._h:
            L.lib.sdb_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def replay(server, requests, arrivals_s):
    """Submits requests [(engine_index, plan)] at the given arrival offsets
    (seconds from now), then waits for all; returns [(Result, info)] in
    request order."""
    import time
    t0 = time.perf_counter()
    tickets = []
    for (e, plan), at in zip(requests, arrivals_s):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(server.submit(e, plan))
    return [server.wait(t) for t in tickets]
