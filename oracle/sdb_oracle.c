/*
 * CPU oracle for the SYCLDB physical-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it restates, in
 * plain C11 + OpenMP, what the reference's device functors compute, one
 * function per functor, so that the HIP kernels of the product can be
 * compared against it bit for bit.  Nothing in the product links or calls
 * it; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it (through oracle/liboracle.so).
 *
 * Every function cites the reference file:line it follows.  Integer
 * semantics are made explicit so that they do not depend on C's undefined
 * signed overflow:
 *   - int32 arithmetic wraps (computed in uint32, reinterpreted);
 *   - int32 -> uint64 conversion is the modular (sign-extending) one that
 *     C performs for `uint64_t x = (int)v`;
 *   - division truncates toward zero; the reference leaves x/0 and
 *     INT_MIN/-1 undefined, this restatement (and the product) define
 *     x/0 = 0 and INT_MIN/-1 = INT_MIN.  These two cases are "parity
 *     unpinned" because the reference has no defined result for them.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* comparison / logical / arithmetic enums: kernels/selection.hpp:7-21,
 * kernels/aggregation.hpp:11-17 (same numeric order). */
enum { ORC_EQ = 0, ORC_NE, ORC_LT, ORC_LE, ORC_GT, ORC_GE };
enum { ORC_NONE = 0, ORC_AND, ORC_OR };
enum { ORC_MUL = 0, ORC_DIV, ORC_ADD, ORC_SUB };

/* compare(): kernels/selection.hpp:61-81 (unknown op -> false) */
static inline int orc_compare(int cmp, int32_t a, int32_t b)
{
    switch (cmp) {
    case ORC_EQ: return a == b;
    case ORC_NE: return a != b;
    case ORC_LT: return a < b;
    case ORC_LE: return a <= b;
    case ORC_GT: return a > b;
    case ORC_GE: return a >= b;
    default: return 0;
    }
}

/* logical(): kernels/selection.hpp:83-96.  NONE returns b (overwrites). */
static inline int orc_logical_op(int logic, int a, int b)
{
    switch (logic) {
    case ORC_AND: return a && b;
    case ORC_OR: return a || b;
    case ORC_NONE: return b;
    default: return 0;
    }
}

/* element_operation(): kernels/aggregation.hpp:19-35, int32 wrapping. */
static inline int32_t orc_element_op(int op, int32_t a, int32_t b)
{
    uint32_t ua = (uint32_t)a, ub = (uint32_t)b;
    switch (op) {
    case ORC_MUL: return (int32_t)(ua * ub);
    case ORC_ADD: return (int32_t)(ua + ub);
    case ORC_SUB: return (int32_t)(ua - ub);
    case ORC_DIV:
        if (b == 0) return 0;
        if (a == INT32_MIN && b == -1) return INT32_MIN;
        return a / b;
    default: return 0;
    }
}

/* HASH(X, Y, Z) = (X - Z) % Y : kernels/join.hpp:11-15.  Callers only use
 * it for X in [Z, Z + Y - 1], where it equals X - Z; the subtraction is
 * done in uint32 so an out-of-range X (whose hash the reference computes
 * and then discards, join.hpp:296) has no undefined behaviour here. */
static inline int32_t orc_hash(int32_t x, int32_t len, int32_t min)
{
    int32_t d = (int32_t)((uint32_t)x - (uint32_t)min);
    return d % len;
}

/* SelectionKernelLiteral::operator(): kernels/selection.hpp:182-185 */
void orc_select_lit(const int32_t *col, const uint8_t *in, uint8_t *out,
                    int64_t n, int cmp, int logic, int32_t v)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = (uint8_t)orc_logical_op(logic, in[i] != 0, orc_compare(cmp, col[i], v));
}

/* SelectionKernelColumns::operator(): kernels/selection.hpp:151-154 */
void orc_select_col(const int32_t *a, const int32_t *b, const uint8_t *in,
                    uint8_t *out, int64_t n, int cmp, int logic)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = (uint8_t)orc_logical_op(logic, in[i] != 0, orc_compare(cmp, a[i], b[i]));
}

/* LogicalKernel::operator(): kernels/selection.hpp:121-124 */
void orc_logical(const uint8_t *in, const uint8_t *local, uint8_t *out,
                 int64_t n, int logic)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = (uint8_t)orc_logical_op(logic, in[i] != 0, local[i] != 0);
}

/* FillKernel::operator(): kernels/projection.hpp:18-21 */
void orc_fill_i32(int32_t *dst, int32_t v, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        dst[i] = v;
}

/* PerformOperationKernelColumns: kernels/aggregation.hpp:65-71
 * (rows whose flag is false keep their previous content) */
void orc_binop_cc(int32_t *r, const int32_t *a, const int32_t *b,
                  const uint8_t *f, int64_t n, int op)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        if (f[i]) r[i] = orc_element_op(op, a[i], b[i]);
}

/* PerformOperationKernelLiteralFirst: kernels/aggregation.hpp:118-124 */
void orc_binop_lc(int32_t *r, int32_t a, const int32_t *b,
                  const uint8_t *f, int64_t n, int op)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        if (f[i]) r[i] = orc_element_op(op, a, b[i]);
}

/* PerformOperationKernelLiteralSecond: kernels/aggregation.hpp:171-177 */
void orc_binop_cl(int32_t *r, const int32_t *a, int32_t b,
                  const uint8_t *f, int64_t n, int op)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        if (f[i]) r[i] = orc_element_op(op, a[i], b);
}

/* BuildKeysHTKernel::operator(): kernels/join.hpp:29-32.  Every build row
 * writes its flag (true or false).  Run sequentially so that a repeated key
 * resolves deterministically (last row wins); the reference races there. */
void orc_build_keys(const int32_t *col, const uint8_t *flags, int64_t n,
                    uint8_t *ht, int32_t len, int32_t min)
{
    for (int64_t i = 0; i < n; i++) {
        int32_t h = orc_hash(col[i], len, min);
        if (h >= 0 && h < len) ht[h] = flags[i] ? 1 : 0;
    }
}

/* BuildKeyValsHTKernel::operator(): kernels/join.hpp:81-92, interleaved
 * [flag, value] int pairs. */
void orc_build_keyvals(const int32_t *col, const int32_t *vals,
                       const uint8_t *flags, int64_t n, int32_t *ht,
                       int32_t len, int32_t min)
{
    for (int64_t i = 0; i < n; i++) {
        int32_t h = orc_hash(col[i], len, min);
        if (h < 0 || h >= len) continue;
        if (flags[i]) {
            ht[2 * (int64_t)h] = 1;
            ht[2 * (int64_t)h + 1] = vals[i];
        } else {
            ht[2 * (int64_t)h] = 0;
        }
    }
}

/* FilterJoinKernel::operator(): kernels/join.hpp:161-172 (semi-join probe;
 * ht_len = max - min + 1, join.hpp:148) */
void orc_probe_semi(const int32_t *probe, const uint8_t *in, uint8_t *out,
                    int64_t n, const uint8_t *ht, int32_t min, int32_t max)
{
    int32_t len = (int32_t)((uint32_t)max - (uint32_t)min + 1u);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int32_t k = probe[i];
        if (in[i] && k >= min && k <= max)
            out[i] = ht[orc_hash(k, len, min)] ? 1 : 0;
        else
            out[i] = 0;
    }
}

/* FullJoinKernel::operator(): kernels/join.hpp:291-311 (ht_len from
 * join.hpp:277).  val_out of a row that is not selected is left as is. */
void orc_probe_full(const int32_t *probe, int32_t *val_out, const uint8_t *in,
                    uint8_t *out, int64_t n, const int32_t *ht, int32_t min,
                    int32_t max)
{
    int32_t len = (int32_t)((uint32_t)max - (uint32_t)min + 1u);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (in[i]) {
            int32_t k = probe[i];
            if (k >= min && k <= max) {
                int64_t h = (int64_t)orc_hash(k, len, min) * 2;
                if (ht[h] == 1) {
                    val_out[i] = ht[h + 1];
                    out[i] = 1;
                    continue;
                }
            }
            out[i] = 0;
        } else {
            out[i] = 0;
        }
    }
}

/* AggregateOperationKernel / aggregate_operation: kernels/aggregation.hpp:
 * 224-227 and 258-264: sum of (uint64)(int32) over flagged rows, added to
 * the existing *acc (no initialize_to_identity, aggregation.hpp:257). */
void orc_sum_flagged(const int32_t *col, const uint8_t *flags, int64_t n,
                     uint64_t *acc)
{
    uint64_t s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; i++)
        if (flags[i]) s += (uint64_t)(int64_t)col[i];
    *acc += s;
}

/* Dense group-by slot of one row: kernels/aggregation.hpp:315-321.  The
 * index is accumulated in int32 (wrapping), then `hash %= prod_ranges`
 * converts it to uint64 first (so a negative index wraps modulo 2^64). */
static inline int64_t orc_group_slot(const int32_t *const *keys, int k,
                                     const int32_t *mins, const int32_t *maxs,
                                     uint64_t prod, int64_t i)
{
    uint32_t hash = 0, mult = 1;
    for (int j = 0; j < k; j++) {
        hash += ((uint32_t)keys[j][i] - (uint32_t)mins[j]) * mult;
        mult *= (uint32_t)maxs[j] - (uint32_t)mins[j] + 1u;
    }
    uint64_t h64 = (uint64_t)(int64_t)(int32_t)hash;
    return (int64_t)(int32_t)(uint32_t)(h64 % prod);
}

/* GroupByAggregateKernel::operator(): kernels/aggregation.hpp:310-343.
 * sums/seen must come zeroed (alloc_zero, transient_table.hpp:1480-1481)
 * and are accumulated into.  The reference lets the first writer of a slot
 * (exchange on seen) store the key values; here the first flagged row in
 * row order does, which is one of the orders the reference allows.
 * Threads own contiguous row ranges and are merged in thread order, so the
 * result equals the sequential loop. */
void orc_groupby_dense(const int32_t *const *keys, int k, const int32_t *mins,
                       const int32_t *maxs, const int32_t *agg,
                       const uint8_t *flags, int64_t n, int32_t *const *out_keys,
                       uint64_t *sums, uint32_t *seen, uint64_t prod)
{
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    /* private copies only when they are small */
    if (nthreads > 1 && prod <= (1u << 20) && n > (1 << 16)) {
        uint64_t *psum = (uint64_t *)calloc((size_t)nthreads * prod, sizeof(uint64_t));
        int64_t *pfirst = (int64_t *)malloc((size_t)nthreads * prod * sizeof(int64_t));
        if (psum && pfirst) {
            for (uint64_t x = 0; x < (uint64_t)nthreads * prod; x++) pfirst[x] = -1;
#pragma omp parallel num_threads(nthreads)
            {
                int t = 0;
#ifdef _OPENMP
                t = omp_get_thread_num();
#endif
                uint64_t *ms = psum + (uint64_t)t * prod;
                int64_t *mf = pfirst + (uint64_t)t * prod;
#pragma omp for schedule(static)
                for (int64_t i = 0; i < n; i++) {
                    if (!flags[i]) continue;
                    int64_t s = orc_group_slot(keys, k, mins, maxs, prod, i);
                    if (mf[s] < 0) mf[s] = i;
                    ms[s] += (uint64_t)(int64_t)agg[i];
                }
            }
#pragma omp parallel for schedule(static)
            for (int64_t s = 0; s < (int64_t)prod; s++) {
                int64_t first = -1;
                uint64_t acc = 0;
                for (int t = 0; t < nthreads; t++) {
                    acc += psum[(uint64_t)t * prod + s];
                    int64_t f = pfirst[(uint64_t)t * prod + s];
                    if (first < 0 && f >= 0) first = f;
                }
                if (first >= 0) {
                    if (seen[s] == 0)
                        for (int j = 0; j < k; j++) out_keys[j][s] = keys[j][first];
                    seen[s] = 1;
                }
                sums[s] += acc;
            }
            free(psum);
            free(pfirst);
            return;
        }
        free(psum);
        free(pfirst);
    }
    for (int64_t i = 0; i < n; i++) {
        if (!flags[i]) continue;
        int64_t s = orc_group_slot(keys, k, mins, maxs, prod, i);
        if (seen[s] == 0) {
            seen[s] = 1;
            for (int j = 0; j < k; j++) out_keys[j][s] = keys[j][i];
        }
        sums[s] += (uint64_t)(int64_t)agg[i];
    }
}

/* final flags of a group-by: aggregation.hpp:517-530,
 * transient_table.hpp:1544-1558 */
void orc_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)prod; i++)
        flags[i] = seen[i] != 0;
}

/* count_true_flags: kernels/common.hpp:30-79 (popcount into uint64) */
uint64_t orc_count_flags(const uint8_t *flags, int64_t n)
{
    uint64_t c = 0;
#pragma omp parallel for schedule(static) reduction(+ : c)
    for (int64_t i = 0; i < n; i++)
        c += flags[i] != 0;
    return c;
}

/* build_row_ids: models/transient_table.hpp:535-599 collects the indices of
 * true flags (in an order the reference leaves unspecified).  The oracle
 * returns them ascending; tests compare as sets or against ascending order
 * for the ordered product kernel. */
int64_t orc_compact(const uint8_t *flags, int64_t n, int32_t *row_ids)
{
    int64_t c = 0;
    for (int64_t i = 0; i < n; i++)
        if (flags[i]) row_ids[c++] = (int32_t)i;
    return c;
}

/* per-segment min/max of a base column: models/models.hpp:71-77 */
void orc_minmax(const int32_t *col, int64_t n, int32_t *mn, int32_t *mx)
{
    int32_t lo = col[0], hi = col[0];
#pragma omp parallel for schedule(static) reduction(min : lo) reduction(max : hi)
    for (int64_t i = 1; i < n; i++) {
        if (col[i] < lo) lo = col[i];
        if (col[i] > hi) hi = col[i];
    }
    *mn = lo;
    *mx = hi;
}

/* ------------------------------------------------------------------------
 * Synthetic star-schema generator (the data spec shared with the product's
 * device generator; not a reference algorithm).  SplitMix64 (Steele, Lea,
 * Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) as a
 * counter-based hash, and Lemire's multiply-shift range reduction
 * ("Fast random integer generation in an interval", TOMACS 2019).
 * --------------------------------------------------------------------- */
static inline uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t orc_stream_key(uint64_t seed, uint64_t stream)
{
    return orc_splitmix64(seed * 0x100000001B3ULL + stream);
}

/* uniform integer in [lo, hi] for row i of a stream */
static inline int32_t orc_uniform(uint64_t key, uint64_t i, int32_t lo, int32_t hi)
{
    uint64_t h = orc_splitmix64(key + i);
    uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
    uint64_t r = ((h >> 32) * range) >> 32;
    return (int32_t)((int64_t)lo + (int64_t)r);
}

/* out[i] = lo + U(seed, stream, first_row + i) for i < n */
void orc_gen_uniform(int32_t *out, int64_t n, uint64_t seed, uint64_t stream,
                     int64_t first_row, int32_t lo, int32_t hi)
{
    uint64_t key = orc_stream_key(seed, stream);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = orc_uniform(key, (uint64_t)(first_row + i), lo, hi);
}

/* out[i] = base * mult + U(...) (the hierarchical nation/city/brand columns) */
void orc_gen_child(int32_t *out, const int32_t *base, int64_t n, uint64_t seed,
                   uint64_t stream, int64_t first_row, int32_t mult, int32_t lo,
                   int32_t hi)
{
    uint64_t key = orc_stream_key(seed, stream);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = base[i] * mult + orc_uniform(key, (uint64_t)(first_row + i), lo, hi);
}

/* out[i] = table[U(0, table_len - 1)] (fact foreign keys into a dimension
 * whose keys are not 1..N, i.e. the yyyymmdd datekeys) */
void orc_gen_lookup(int32_t *out, int64_t n, uint64_t seed, uint64_t stream,
                    int64_t first_row, const int32_t *table, int32_t table_len)
{
    uint64_t key = orc_stream_key(seed, stream);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = table[orc_uniform(key, (uint64_t)(first_row + i), 0, table_len - 1)];
}

/* days-from-civil inverse: H. Hinnant, "chrono-Compatible Low-Level Date
 * Algorithms" (civil_from_days).  Fills the date dimension: datekey
 * yyyymmdd, year, yearmonthnum for `ndays` days starting 1992-01-01. */
void orc_gen_dates(int32_t *datekey, int32_t *year, int32_t *yearmonthnum,
                   int32_t ndays)
{
    /* days since 1970-01-01 of 1992-01-01 */
    const int64_t start = 8035;
    for (int32_t d = 0; d < ndays; d++) {
        int64_t z = start + d + 719468;
        int64_t era = (z >= 0 ? z : z - 146096) / 146097;
        int64_t doe = z - era * 146097;
        int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
        int64_t y = yoe + era * 400;
        int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
        int64_t mp = (5 * doy + 2) / 153;
        int64_t dd = doy - (153 * mp + 2) / 5 + 1;
        int64_t m = mp < 10 ? mp + 3 : mp - 9;
        y += (m <= 2);
        datekey[d] = (int32_t)(y * 10000 + m * 100 + dd);
        year[d] = (int32_t)y;
        yearmonthnum[d] = (int32_t)(y * 100 + m);
    }
}

/* Host memory bandwidth next to the CPU baseline (not part of the
 * restatement): the best of five parallel copies of `bytes` bytes, counted
 * as read + write, in GB/s (a STREAM "copy"). */
double orc_host_copy_gbps(int64_t bytes)
{
    const int64_t n = bytes / 8;
    if (n <= 0) return 0.0;
    uint64_t *a = (uint64_t *)malloc((size_t)n * 8), *b = (uint64_t *)malloc((size_t)n * 8);
    if (!a || !b) {
        free(a);
        free(b);
        return 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        a[i] = (uint64_t)i;
        b[i] = 0;
    }
    double best = 0.0;
    for (int r = 0; r < 5; r++) {
        const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) b[i] = a[i];
        const double t = omp_get_wtime() - t0;
        if (t > 0 && 2.0 * (double)n * 8 / t / 1e9 > best) best = 2.0 * (double)n * 8 / t / 1e9;
    }
    volatile uint64_t sink = b[n / 2];
    (void)sink;
    free(a);
    free(b);
    return best;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
