/*
 * sycldb.h -- C ABI of the MI355X-native SYCLDB physical-operator library.
 *
 * Drop-in boundary for the operator hot path of SYCLDB (reference files
 * kernels/selection.hpp, kernels/projection.hpp, kernels/join.hpp,
 * kernels/aggregation.hpp, kernels/common.hpp) and for the executor that
 * drives them (models/transient_table.hpp, executor/ddor_executor.cpp).
 *
 * Conventions (replacing the reference's sycl::queue / sycl::event):
 *   - every pointer argument of an operator is a device pointer unless the
 *     comment says "host";
 *   - columns are int32, selection flags are 1-byte booleans holding 0 or 1
 *     (sizeof(bool) == 1 in the reference), sums are uint64 holding the
 *     two's-complement int64 sum;
 *   - `stream` is a hipStream_t (NULL = the default stream); operators are
 *     asynchronous and ordered by the stream, as the reference's kernels are
 *     ordered by event dependencies;
 *   - lengths are int64 (the reference's `int col_len` is at most 2^28);
 *   - the return value is a status: SDB_OK, or an SDB_ERR_* code whose text
 *     sdb_last_error() returns (per thread).  No C++ exception crosses the
 *     ABI.  Where the reference throws std::invalid_argument (unknown
 *     arithmetic op) this returns SDB_ERR_INVALID_ARG; where it throws
 *     std::bad_alloc this returns SDB_ERR_OOM.
 */
#ifndef SYCLDB_H
#define SYCLDB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *sdb_stream; /* hipStream_t */

enum sdb_status {
    SDB_OK = 0,
    SDB_ERR_INVALID_ARG = 1,
    SDB_ERR_HIP = 2,
    SDB_ERR_OOM = 3,
    SDB_ERR_UNSUPPORTED = 4,
    SDB_ERR_RUNTIME = 5
};

/* comp_op, kernels/selection.hpp:7-15 (same values) */
enum sdb_comp_op { SDB_EQ = 0, SDB_NE, SDB_LT, SDB_LE, SDB_GT, SDB_GE };
/* logical_op, kernels/selection.hpp:16-21 */
enum sdb_logical_op { SDB_NONE = 0, SDB_AND, SDB_OR };
/* BinaryOp, kernels/aggregation.hpp:11-17 */
enum sdb_binop { SDB_MUL = 0, SDB_DIV, SDB_ADD, SDB_SUB };

/* Text of the last error on this thread ("" if none). */
const char *sdb_last_error(void);
/* Library version string. */
const char *sdb_version(void);

/* get_comp_op, kernels/selection.hpp:23-39: unknown strings map to SDB_EQ. */
int sdb_comp_op_from_string(const char *op);
/* get_logical_op, kernels/selection.hpp:41-49: unknown / "" map to SDB_NONE. */
int sdb_logical_op_from_string(const char *op);
/* get_op_from_string, kernels/aggregation.hpp:37-48: -1 for an unknown op
 * (the reference throws std::invalid_argument). */
int sdb_binop_from_string(const char *op);

/* ------------------------------------------------------------ selection */

/* SelectionKernelLiteral, kernels/selection.hpp:157-186, and
 * selection(flags, arr, op, value, parent_op, ...) 188-214:
 *   out[i] = logical(logic, in[i], compare(cmp, col[i], value)).
 * in == out is allowed (the single-pointer constructor, 170-172). */
int sdb_select_lit(const int32_t *col, const uint8_t *in_flags, uint8_t *out_flags, int64_t n,
                   int cmp, int logic, int32_t value, sdb_stream stream);

/* SelectionKernelColumns, kernels/selection.hpp:127-155, and
 * selection(flags, op1, op, op2, parent_op, ...) 216-242. */
int sdb_select_col(const int32_t *a, const int32_t *b, const uint8_t *in_flags, uint8_t *out_flags,
                   int64_t n, int cmp, int logic, sdb_stream stream);

/* LogicalKernel, kernels/selection.hpp:106-125:
 *   out[i] = logical(logic, in[i], local[i]). */
int sdb_logical(const uint8_t *in_flags, const uint8_t *local_flags, uint8_t *out_flags, int64_t n,
                int logic, sdb_stream stream);

/* ----------------------------------------------------------- projection */

/* FillKernel, kernels/projection.hpp:7-21: dst[i] = value. */
int sdb_fill_i32(int32_t *dst, int32_t value, int64_t n, sdb_stream stream);

/* PerformOperationKernelColumns + perform_operation(result, a[], b[], ...),
 * kernels/aggregation.hpp:50-100: if (flags[i]) result[i] = a[i] op b[i]
 * (int32, wrapping).  Rows whose flag is 0 are not written. */
int sdb_binop_cc(int32_t *result, const int32_t *a, const int32_t *b, const uint8_t *flags,
                 int64_t n, int op, sdb_stream stream);
/* PerformOperationKernelLiteralFirst, kernels/aggregation.hpp:102-153:
 * result[i] = a op b[i]. */
int sdb_binop_lc(int32_t *result, int32_t a, const int32_t *b, const uint8_t *flags, int64_t n,
                 int op, sdb_stream stream);
/* PerformOperationKernelLiteralSecond, kernels/aggregation.hpp:155-206:
 * result[i] = a[i] op b. */
int sdb_binop_cl(int32_t *result, const int32_t *a, int32_t b, const uint8_t *flags, int64_t n,
                 int op, sdb_stream stream);

/* ----------------------------------------------------------------- join */

/* BuildKeysHTKernel / build_keys_ht, kernels/join.hpp:17-58:
 *   ht[key - ht_min] = flags[i] for every row (uint8 table of ht_len slots).
 * Rows whose key falls outside [ht_min, ht_min + ht_len) are skipped (the
 * reference writes out of bounds there). */
int sdb_build_keys(const int32_t *keys, const uint8_t *flags, int64_t n, uint8_t *ht,
                   int32_t ht_len, int32_t ht_min, sdb_stream stream);

/* BuildKeyValsHTKernel / build_key_vals_ht, kernels/join.hpp:60-127:
 * interleaved int32 pairs ht[2h] = flag ? 1 : 0, ht[2h+1] = vals[i] when the
 * flag is set (the value slot is left alone otherwise). */
int sdb_build_keyvals(const int32_t *keys, const int32_t *vals, const uint8_t *flags, int64_t n,
                      int32_t *ht, int32_t ht_len, int32_t ht_min, sdb_stream stream);

/* FilterJoinKernel / filter_join, kernels/join.hpp:129-208 (semi join):
 *   out[i] = in[i] && min <= probe[i] <= max && ht[probe[i] - min]. */
int sdb_probe_semi(const int32_t *probe, const uint8_t *in_flags, uint8_t *out_flags, int64_t n,
                   const uint8_t *ht, int32_t ht_min, int32_t ht_max, sdb_stream stream);

/* FullJoinKernel / full_join, kernels/join.hpp:256-355: when in[i] and the
 * key is in range and ht[2h] == 1: val_out[i] = ht[2h+1], out[i] = 1;
 * otherwise out[i] = 0 and val_out[i] is not written. */
int sdb_probe_full(const int32_t *probe, int32_t *val_out, const uint8_t *in_flags,
                   uint8_t *out_flags, int64_t n, const int32_t *ht, int32_t ht_min,
                   int32_t ht_max, sdb_stream stream);

/* ---------------------------------------------------------- aggregation */

/* AggregateOperationKernel / aggregate_operation,
 * kernels/aggregation.hpp:208-276: *acc += sum over flagged rows of
 * (uint64)(int64)col[i].  *acc is NOT reset (the reference reduction has no
 * initialize_to_identity). */
int sdb_sum_flagged(const int32_t *col, const uint8_t *flags, int64_t n, uint64_t *acc,
                    sdb_stream stream);

/* GroupByAggregateKernel / group_by_aggregate (pointer API),
 * kernels/aggregation.hpp:278-409.  `keys` and `out_keys` are HOST arrays of
 * k (1..8) device pointers; `mins`/`maxs` are HOST arrays of k ints.  For
 * each flagged row: slot = (sum_j (keys[j][i]-mins[j]) * prod_{l<j}
 * (maxs[l]-mins[l]+1)) (int32 arithmetic) % prod; seen[slot] = 1 (the first
 * writer stores the key values into out_keys[j][slot]); sums[slot] +=
 * (uint64)(int64)agg[i].  sums/seen are accumulated into (come zeroed). */
int sdb_groupby_dense(const int32_t *const *keys, int k, const int32_t *mins, const int32_t *maxs,
                      const int32_t *agg, const uint8_t *flags, int64_t n, int32_t *const *out_keys,
                      uint64_t *sums, uint32_t *seen, uint64_t prod, sdb_stream stream);

/* Final flags of a group-by, kernels/aggregation.hpp:517-530 and
 * models/transient_table.hpp:1544-1558: flags[g] = seen[g] != 0. */
int sdb_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod, sdb_stream stream);

/* count_true_flags (kernels/common.hpp:30-110): *out = number of set flags
 * when accumulate == 0 (initialize_to_identity, common.hpp:317), or
 * *out += that number when accumulate != 0 (the allocator overload). */
int sdb_count_flags(const uint8_t *flags, int64_t n, uint64_t *out, int accumulate,
                    sdb_stream stream);

/* Stream compaction of flags into row ids, the device half of
 * build_row_ids (models/transient_table.hpp:535-599).  The reference's order
 * is unspecified; this one writes the ids ascending.  *count (device) gets
 * the number of ids; row_ids needs room for n entries. */
int sdb_compact(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count,
                sdb_stream stream);

/* sdb_compact for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same results as sdb_compact; it
 * allocates nothing: its per-tile counts live in a fixed 256 KiB device array
 * of the library (one per GPU it runs on, held until the process ends), so n
 * is at most SDB_COMPACT_NOALLOC_MAX_ROWS (2^28, the reference's
 * SEGMENT_SIZE); a larger n returns SDB_ERR_INVALID_ARG.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (on different streams
 * or threads) race. */
#define SDB_COMPACT_NOALLOC_MAX_ROWS (1LL << 28)
int sdb_compact_noalloc(const uint8_t *flags, int64_t n, int32_t *row_ids, uint32_t *count,
                        sdb_stream stream);

/* Gather of selected rows, the device half of Segment::compress_sync
 * (models/models.hpp:1019-1035): dst[j] = src[row_ids[j]] for j < count
 * (count read on the host by the caller). */
int sdb_gather_i32(const int32_t *src, const int32_t *row_ids, int64_t count, int32_t *dst,
                   sdb_stream stream);

/* Segment::compress_sync driven by TransientTable::build_row_ids
 * (models/models.hpp:990-1068, models/transient_table.hpp:535-599), the
 * sync of a device segment into its host copy: for every row whose device
 * flag is set, host_cols[c][row] = cols[c][row], for the ncols (0..8)
 * device columns of n rows; the host columns' other rows are left as they
 * are.  The selected rows are compacted and gathered on the device, copied
 * to the host, and scattered there.  *count (host) receives the number of
 * rows synced.  `cols` and `host_cols` are HOST arrays of pointers; the
 * host columns hold n int32 each.  Synchronises the stream. */
int sdb_compress_sync(const int32_t *const *cols, int ncols, const uint8_t *flags, int64_t n,
                      int32_t *const *host_cols, uint64_t *count, sdb_stream stream);

/* Per-column min and max (the load-time statistics of
 * models/models.hpp:71-77), written to the host ints min_out and max_out.
 * Synchronises the stream.  n must be > 0. */
int sdb_minmax_i32(const int32_t *col, int64_t n, int32_t *min_out, int32_t *max_out,
                   sdb_stream stream);

/* ------------------------------------------------------- fused bundle */

/* Operator kinds of a bundle: KernelType, models/execution.hpp:30-45 (same
 * values). */
enum sdb_op_kind {
    SDB_OP_EMPTY = 0,
    SDB_OP_LOGICAL,       /* LogicalKernel: out = logical(logic, in, local) */
    SDB_OP_SELECT_COL,    /* SelectionKernelColumns: out = logical(logic, in, cmp(col, col2)) */
    SDB_OP_SELECT_LIT,    /* SelectionKernelLiteral: out = logical(logic, in, cmp(col, value)) */
    SDB_OP_FILL,          /* FillKernel: res = value */
    SDB_OP_BINOP_CC,      /* PerformOperationKernelColumns: where in: res = col binop col2 */
    SDB_OP_BINOP_LC,      /* PerformOperationKernelLiteralFirst: where in: res = value binop col */
    SDB_OP_BINOP_CL,      /* PerformOperationKernelLiteralSecond: where in: res = col binop value */
    SDB_OP_BUILD_KEYS,    /* BuildKeysHTKernel: ht (uint8, ht_len slots from ht_min)[col] = in */
    SDB_OP_PROBE_SEMI,    /* FilterJoinKernel: out = in && ht_min <= col <= ht_max && ht[col - ht_min] */
    SDB_OP_BUILD_KEYVALS, /* BuildKeyValsHTKernel: ht (int32 pairs) from keys col, values col2, flags in */
    SDB_OP_PROBE_FULL,    /* FullJoinKernel: match -> res = value, out = 1; else out = 0 */
    SDB_OP_SUM,           /* AggregateOperationKernel: *acc += sum of (uint64)(int64)col where in */
    SDB_OP_GROUPBY        /* GroupByAggregateKernel: as sdb_groupby_dense, aggregating col where in */
};
#define SDB_OP_MAX_KEYS 8
#define SDB_BUNDLE_MAX_OPS 64

/* One operator of a bundle: the arguments of one KernelData
 * (models/execution.hpp:30-497).  The first fields are the reference's
 * SYCLDBContext slot (kernels/jit_kernels.hpp:12-20): col_ptrs -> col,
 * col_ptrs2 -> col2, res_ptrs -> res, values -> value, ht_ptrs and
 * ht_int_ptrs -> ht, params[i*8 + 0..5] -> cmp, logic, binop, ht_min, ht_max,
 * ht_len.  The rest is what the slot lacks: the operator's own flag buffers
 * (the reference threads one pass flag through the fused prefix instead),
 * the SUM accumulator and the group-by arguments.  Fields an operator kind
 * does not use are ignored.  All pointers are device pointers, except that
 * keys / out_keys hold device pointers in the struct itself. */
typedef struct sdb_op {
    int32_t kind;   /* sdb_op_kind */
    int32_t cmp;    /* sdb_comp_op */
    int32_t logic;  /* sdb_logical_op (the parent op of a selection) */
    int32_t binop;  /* sdb_binop */
    int32_t ht_min, ht_max, ht_len; /* probes: ht_len 0 means ht_max - ht_min + 1;
                                       builds: ht_len slots from ht_min */
    int32_t value;  /* literal */
    const int32_t *col, *col2;
    int32_t *res;
    void *ht;
    const uint8_t *in_flags;
    uint8_t *out_flags;
    const uint8_t *local_flags;
    uint64_t *acc;
    int32_t nkeys;
    int32_t mins[SDB_OP_MAX_KEYS], maxs[SDB_OP_MAX_KEYS];
    const int32_t *keys[SDB_OP_MAX_KEYS];
    int32_t *out_keys[SDB_OP_MAX_KEYS];
    uint64_t *sums;
    uint32_t *seen;
    uint64_t prod;
} sdb_op;

/* KernelBundle::execute (models/execution.hpp:528-650) with its fused path
 * (546-615, switched off in the reference) switched on: the nops operators,
 * all over rows [0, n) of one segment, run in order on `stream` with the
 * results of running them one by one -- as one fused kernel when the
 * operators allow it (every buffer an operator writes is written back),
 * otherwise one kernel per operator.  The caller owns every buffer; each is
 * checked against the size its operator implies (n rows, ht_len slots, prod
 * slots).  n <= 2^31 - 1, nops <= SDB_BUNDLE_MAX_OPS. */
int sdb_bundle_exec(const sdb_op *ops, int nops, int64_t n, sdb_stream stream);

/* ------------------------------------------- device memory and streams */

/* Conveniences for callers without the HIP headers (an FFI, the C++ layer
 * in sycldb.hpp); the reference's own allocations come from its
 * memory_manager / sycl::malloc_* instead.  Copies are ordered on `stream`
 * and complete before the call returns. */
int sdb_device_alloc(void **ptr, uint64_t bytes);
int sdb_device_free(void *ptr);
int sdb_copy_to_device(void *dst, const void *host_src, uint64_t bytes, sdb_stream stream);
int sdb_copy_to_host(void *host_dst, const void *src, uint64_t bytes, sdb_stream stream);
/* Waits for the work queued on `stream` (sycl::event::wait_and_throw). */
int sdb_stream_synchronize(sdb_stream stream);

/* ------------------------------------------------ synthetic data (seeded) */

/* out[i] = lo + U(seed, stream_id, first_row + i) over [lo, hi]
 * (SplitMix64 counter hash + multiply-shift range reduction). */
int sdb_gen_uniform(int32_t *out, int64_t n, uint64_t seed, uint64_t stream_id, int64_t first_row,
                    int32_t lo, int32_t hi, sdb_stream stream);
/* out[i] = base[i] * mult + U(seed, stream_id, first_row + i) over [lo, hi]. */
int sdb_gen_child(int32_t *out, const int32_t *base, int64_t n, uint64_t seed, uint64_t stream_id,
                  int64_t first_row, int32_t mult, int32_t lo, int32_t hi, sdb_stream stream);
/* out[i] = table[U(seed, stream_id, first_row + i) over [0, table_len-1]]. */
int sdb_gen_lookup(int32_t *out, int64_t n, uint64_t seed, uint64_t stream_id, int64_t first_row,
                   const int32_t *table, int32_t table_len, sdb_stream stream);
/* out[i] = first_row + i + 1. */
int sdb_gen_rowid(int32_t *out, int64_t n, int64_t first_row, sdb_stream stream);

/* -------------------------------------------------------------- engine */

/* The executor side: tables resident in HBM, plans run through the
 * TransientTable / KernelBundle model (models/transient_table.hpp,
 * models/execution.hpp) driven by execute_ddor_plan
 * (executor/ddor_executor.cpp:22-245). */
typedef struct sdb_engine sdb_engine;
typedef struct sdb_result sdb_result;

/* Engine options (all 0 = default). */
typedef struct sdb_engine_options {
    uint64_t device_budget_bytes; /* scratch budget of the device allocator
                                     (memory_manager; default 8 GiB,
                                     common.hpp:7) */
    uint64_t segment_size;        /* rows per segment (default 2^30; the
                                     reference's SEGMENT_SIZE is 2^28,
                                     common.hpp:12; results do not depend on
                                     it) */
    int disable_fusion;           /* 0 (default): each bundle runs as one
                                     fused kernel; 1: one kernel per operator,
                                     as the reference runs with
                                     disable_fusion = true
                                     (models/execution.hpp:546) */
    int device;                   /* HIP device ordinal */
} sdb_engine_options;

int sdb_engine_create(const sdb_engine_options *opts, sdb_engine **out);
void sdb_engine_destroy(sdb_engine *eng);
/* Stream the engine launches on (hipStream_t). */
sdb_stream sdb_engine_stream(sdb_engine *eng);

/* Registers table `name` (one of the SSB tables) with `nrows` rows.  Column
 * storage is allocated per column by sdb_engine_column_alloc. */
int sdb_engine_table_create(sdb_engine *eng, const char *name, uint64_t nrows);
/* Device buffer of column `col` of table `name` (allocated on first call,
 * nrows int32).  The caller fills it (generator, copy). */
int sdb_engine_column_alloc(sdb_engine *eng, const char *name, int col, int32_t **dev_ptr);
/* Copies column `col` (nrows int32) into host memory. */
int sdb_engine_column_download(sdb_engine *eng, const char *name, int col, int32_t *host);
/* Copies host data (nrows int32) into column `col`. */
int sdb_engine_column_upload(sdb_engine *eng, const char *name, int col, const int32_t *host);
/* Loads the columns of `name` listed by table_column_indices
 * (kernels/types.cpp:3-11) from `<dir>/<TABLE><col>` raw int32 files
 * (models/models.hpp:1659-1708). */
int sdb_engine_table_load_dir(sdb_engine *eng, const char *name, const char *dir);
/* Computes per-segment min/max of every column of the table (the Segment
 * constructor statistics); call after the columns are filled. */
int sdb_engine_table_finalize(sdb_engine *eng, const char *name);

/* Runs a plan (JSON mirror of calciteserver.thrift PlanResult) to
 * completion and returns its result table. */
int sdb_engine_execute(sdb_engine *eng, const char *plan_json, sdb_result **out);
/* Runs n plans as one batch (no counterpart in the reference): the
 * dimension-side work of every plan runs as in sdb_engine_execute; the
 * fact-table bundles of the plans that fuse run together in one kernel per
 * segment that scans the union of their fact columns once; the other plans
 * run alone afterwards.  out[i] receives plan i's result, identical to
 * running it alone; *batched (may be NULL) = how many shared the scan. */
int sdb_engine_execute_batch(sdb_engine *eng, const char *const *plans, int n, sdb_result **out, int *batched);

/* ----------------------------------------------------------- serving */

/* A serving loop over engines with resident tables (no counterpart in the
 * reference, which runs one query at a time).  Requests are plans for one
 * engine, queued per engine with their arrival time.  With batching
 * (serial == 0) each engine has its own worker (single_worker: one worker
 * for all, taking the engine whose oldest waiting request arrived first), so
 * the engines' host work and kernels overlap on the device.  When enough
 * requests wait for an engine for sharing its scan to pay (a threshold
 * learned per engine from the plain and batched run times it has seen), its
 * worker admits them in arrival order while a slot (max_batch) and the
 * memory budget (the requests' estimated build-table and group-by
 * footprints) allow -- always at least one -- and runs them as one batch
 * sharing the fact scan (sdb_engine_execute_batch); otherwise it runs the
 * oldest request alone.  Requests arriving meanwhile join a later batch;
 * each result is handed to its waiter as soon as it is ready.  With
 * serial != 0 one worker runs every request alone in arrival order (the
 * sequential baseline).  While a server runs, its engines must not be used
 * directly. */
typedef struct sdb_server sdb_server;
typedef struct sdb_server_options {
    int max_batch;                /* slots per batch (default 32) */
    uint64_t memory_budget_bytes; /* admission budget per batch (default 2 GiB) */
    int serial;                   /* 1: one request at a time, arrival order */
    int single_worker;            /* 1: one worker for every engine (default:
                                     one per engine, so that the engines'
                                     host work and kernels overlap) */
} sdb_server_options;
typedef struct sdb_request_info {
    double latency_ms; /* completion - arrival */
    double queue_ms;   /* start - arrival */
    double run_ms;     /* completion - start */
    int batch_size;    /* requests in the batch that ran it (1: alone) */
} sdb_request_info;

int sdb_server_create(sdb_engine *const *engines, int n_engines, const sdb_server_options *opts,
                      sdb_server **out);
/* Queues plan_json for engine index `engine`; *ticket identifies it. */
int sdb_server_submit(sdb_server *srv, int engine, const char *plan_json, int64_t *ticket);
/* Waits for the request and returns its result (and timing, if info). */
int sdb_server_wait(sdb_server *srv, int64_t ticket, sdb_result **out, sdb_request_info *info);
/* Batches run so far and the requests they held. */
int sdb_server_stats(sdb_server *srv, int64_t *batches, int64_t *requests);
/* Finishes the queued requests, then stops the workers. */
void sdb_server_destroy(sdb_server *srv);

/* Batched launches (no counterpart in the reference): the programs and
 * mask descriptors of a launch reach the device in one copy, then the mask
 * build and the batched kernel run.  With graphs on (1), a launch met a
 * second time (same geometry and kernel arguments) is captured into a
 * hipGraph and replayed for later such launches, from device and pinned
 * host buffers kept per shape library; 0 (the default unless
 * SDB_BATCH_GRAPH=1) issues the three operations on the stream each time;
 * -1 restores the default.  With graphs on, batches over one fact table must
 * not run concurrently from several threads. */
int sdb_set_batch_graphs(int on);
/* Graph captures and replays of batched launches so far (either may be NULL). */
int sdb_batch_graph_stats(int64_t *captures, int64_t *replays);

/* Device time (ms, hipEvents) and host wall time (ms) of the last execute. */
int sdb_engine_last_timing(sdb_engine *eng, double *device_ms, double *wall_ms);
/* Resets the scratch allocator (memory_manager::reset, re-zeroing the
 * zero region), as the reference does between repetitions
 * (runtime/runtime.cpp:331-345). */
int sdb_engine_reset(sdb_engine *eng);
/* Number of kernels launched by the last execute. */
int sdb_engine_last_launches(sdb_engine *eng, int64_t *launches);

/* Switches bundle fusion on (1) or off (0) for later executions. */
int sdb_engine_set_fusion(sdb_engine *eng, int on);

/* How a fused bundle's kernel is produced (no counterpart in the reference,
 * whose JIT fusion is switched off: models/execution.hpp:546-615).
 *   SDB_CODEGEN_AUTO       compiled at run time for the program's shape when
 *                          the hipRTC library loads, else interpreted
 *   SDB_CODEGEN_INTERPRET  the compiled-in program interpreter
 *   SDB_CODEGEN_COMPILE    compiled at run time; an error when unavailable
 * `unroll` = 4-row chunks per thread per loop iteration of a compiled kernel
 * (1, 2 or 4), or 0: the autotuner times the candidate launch configurations
 * (chunks per thread x blocks per CU) on the first large launches of each
 * program shape and keeps the fastest, then times that one once more with
 * its row tail queued (below) and keeps the faster form.  1, 2 or 4 may be
 * combined with SDB_UNROLL_QUEUED: the instructions after a program's last
 * probe through global memory then run from a per-wave queue in LDS that is
 * drained 64 rows at a time, where the program has such a tail (same
 * results). */
enum sdb_codegen { SDB_CODEGEN_AUTO = 0, SDB_CODEGEN_INTERPRET = 1, SDB_CODEGEN_COMPILE = 2 };
#define SDB_UNROLL_QUEUED 16
/* with an explicit unroll: blocks of 512 or 1024 threads instead of 256 (the
 * autotuner's other candidate launch shapes) */
#define SDB_UNROLL_BLOCK512 256
#define SDB_UNROLL_BLOCK1024 512
int sdb_engine_set_codegen(sdb_engine *eng, int mode, int unroll);
/* 1 when runtime compilation is available, else 0 with the reason in `why`
 * (may be NULL). */
int sdb_codegen_available(char *why, int why_cap);
/* Distinct fused kernels compiled at run time in this process. */
int sdb_codegen_kernels_compiled(void);
/* Autotuner state, one text line per program shape; returns the length
 * (the text is truncated to cap - 1 bytes). */
int sdb_codegen_tuning_report(char *buf, int cap);
/* Kernels are compiled on a background thread: in SDB_CODEGEN_AUTO mode a
 * program shape runs on the interpreter kernel until its compiled kernel is
 * ready, and the autotuner's candidates compile meanwhile (SDB_CODEGEN_COMPILE
 * waits for them instead).  sdb_codegen_wait waits until every kernel queued
 * so far is compiled (1) or timeout_s seconds have passed (0).
 * sdb_codegen_pending counts the program shapes still compiling or, having
 * run over at least 4M rows, not yet settled on a launch configuration, plus
 * the batched-kernel libraries still tuning; 0 once every launch seen so far
 * runs its final configuration. */
int sdb_codegen_wait(double timeout_s);
int sdb_codegen_pending(void);
/* Stops the background compiler: queued compilations are dropped (their
 * program shapes run on the interpreter kernel in SDB_CODEGEN_AUTO mode, or
 * fail in SDB_CODEGEN_COMPILE mode) and the one in progress is waited for.
 * Call it before the process exits while kernels may still be compiling: a
 * compilation still running when the process tears down its static objects
 * can crash or hang the exit.  The Python package calls it when the
 * interpreter exits.  No counterpart in the reference (its kernels are
 * compiled ahead of time). */
int sdb_codegen_shutdown(void);
/* Compiles, without launching (no device needed), the runtime-generated
 * kernels of programs covering every instruction form for GPU architecture
 * `arch` (e.g. "gfx950"); *count = kernels compiled. */
int sdb_codegen_check(const char *arch, int *count);

/* Per-launch timing: with profiling on, every kernel launch of a bundle is
 * bracketed by HIP events on the engine stream and accumulated per name
 * ("fused:<table>" for a fused bundle, "op:<KernelType>:<table>" for an
 * operator launched alone) -- the counterpart of the reference's
 * add_profiling_info (models/execution.hpp:18-28). */
int sdb_engine_set_profiling(sdb_engine *eng, int on);
/* Restricts the timing to the bundles of table `table` (NULL or "": every
 * bundle); the others then run without events around their launches. */
int sdb_engine_set_profile_table(sdb_engine *eng, const char *table);
int sdb_engine_profile_reset(sdb_engine *eng);
int sdb_engine_profile_count(sdb_engine *eng);
/* Entry idx (0..count-1): name, launches, rows processed, summed ms. */
int sdb_engine_profile_get(sdb_engine *eng, int idx, char *name, int name_cap, int64_t *launches,
                           int64_t *rows, double *ms);

/* Result table: nrows rows, ncols non-null columns, row selection flags
 * (the final TransientTable's host view, TransientTable::materialize_host_view
 * as executor/result_writer.cpp:28-40 calls it).  Only selected rows carry
 * meaning: the key columns of a group-by table are unspecified at
 * unselected slots (the reference never writes them); large group-by tables
 * are compacted on the device and read 0 there. */
uint64_t sdb_result_nrows(const sdb_result *r);
int sdb_result_ncols(const sdb_result *r);
/* 1 when column c holds uint64 sums, 0 when it holds int32 values. */
int sdb_result_col_is_u64(const sdb_result *r, int c);
/* Copies column c (nrows values, int32 or uint64) into host memory. */
int sdb_result_col_copy(const sdb_result *r, int c, void *host);
/* Copies the nrows selection flags into host memory. */
int sdb_result_flags_copy(const sdb_result *r, uint8_t *host);
/* Number of selected rows. */
uint64_t sdb_result_selected(const sdb_result *r);
/* The selected rows only, as result_writer.cpp:41-67 visits them: their
 * row numbers in ascending order (sdb_result_selected values) and column c's
 * values at those rows (sdb_result_selected values, int32 or uint64). */
int sdb_result_selected_rows(const sdb_result *r, uint64_t *host);
int sdb_result_col_copy_selected(const sdb_result *r, int c, void *host);
/* Writes the .res text of executor/result_writer.cpp:41-67 (selected rows,
 * space-separated, int32 keys and uint64 sums) to `path`. */
int sdb_result_write(const sdb_result *r, const char *path);
void sdb_result_destroy(sdb_result *r);

/* Execution order of a plan (dag_topological_sort,
 * operations/preprocessing.cpp:14-102): writes up to cap rel ids into
 * order[] (host) and their count into *count.  Host-only, no device use. */
int sdb_plan_dag_order(const char *plan_json, int32_t *order, int cap, int *count);

#ifdef __cplusplus
}
#endif

#endif /* SYCLDB_H */
