// Host side of the MI355X SYCLDB engine: the reference's data model and
// dependency model (models/models.hpp, models/execution.hpp,
// models/transient_table.hpp, operations/memory_manager.*) on one HIP
// device, with HIP streams instead of SYCL queues/events.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "fused.hpp"
#include "plan.hpp"
#include "sdb_common.hpp"

namespace sdb {

// ---------------------------------------------------------------- launchers
// (ops.hip)
void launch_select_lit(const int32_t *col, const uint8_t *in, uint8_t *out, int64_t n, int cmp, int logic,
                       int32_t v, hipStream_t s);
void launch_select_col(const int32_t *a, const int32_t *b, const uint8_t *in, uint8_t *out, int64_t n, int cmp,
                       int logic, hipStream_t s);
void launch_logical(const uint8_t *in, const uint8_t *local, uint8_t *out, int64_t n, int logic, hipStream_t s);
void launch_fill_i32(int32_t *dst, int32_t v, int64_t n, hipStream_t s);
void launch_fill_u8(uint8_t *dst, uint8_t v, int64_t n, hipStream_t s);
void launch_binop(int kind, int op, int32_t *r, const int32_t *a, const int32_t *b, int32_t lit, const uint8_t *f,
                  int64_t n, hipStream_t s);
void launch_build_keys(const int32_t *keys, const uint8_t *flags, int64_t n, uint8_t *ht, int32_t len, int32_t min,
                       hipStream_t s);
void launch_build_keyvals(const int32_t *keys, const int32_t *vals, const uint8_t *flags, int64_t n, int32_t *ht,
                          int32_t len, int32_t min, hipStream_t s);
void launch_probe_semi(const int32_t *probe, const uint8_t *in, uint8_t *out, int64_t n, const uint8_t *ht,
                       int32_t min, int32_t max, hipStream_t s);
void launch_probe_full(const int32_t *probe, int32_t *val_out, const uint8_t *in, uint8_t *out, int64_t n,
                       const int32_t *ht, int32_t min, int32_t max, hipStream_t s);
void launch_sum_flagged(const int32_t *col, const uint8_t *flags, int64_t n, uint64_t *acc, hipStream_t s);
void launch_count_flags(const uint8_t *flags, int64_t n, uint64_t *out, bool accumulate, hipStream_t s);
void launch_groupby_dense(const int32_t *const *keys, int k, const int32_t *mins, const int32_t *maxs,
                          const int32_t *agg, const uint8_t *flags, int64_t n, int32_t *const *out_keys,
                          uint64_t *sums, uint32_t *seen, uint64_t prod, hipStream_t s, int strategy);
void launch_seen_to_flags(const uint32_t *seen, uint8_t *flags, uint64_t prod, hipStream_t s);
void minmax_i32(const int32_t *col, int64_t n, int32_t *mn, int32_t *mx, hipStream_t s);
uint64_t groupby_lds_max_slots();
// Compact probe tables (ops.hip): bitmap from a uint8 key table, bitmap and
// uint16 value offsets from a key-value pair table; the set bits counted per
// block into partials[0, bits_partials(len)) (launch_stats_reduce sums
// them), *bad set when a value does not fit 16 bits.
uint32_t bits_partials(int32_t len);
void launch_bits_from_keys(const uint8_t *ht, int32_t len, uint32_t *bits, uint32_t *partials, hipStream_t s);
// Rank dictionary (FV_RANK16 layout) of the presence bitmap of `len` slots
// and their per-slot 16-bit values; `partials` (or null): the per-block
// counts launch_bits_from_pairs left for the same table.
void launch_rank_dict(const uint32_t *bits, int32_t len, const uint16_t *vals, const uint32_t *partials,
                      uint8_t *dict, uint32_t dict_cap, hipStream_t s);
void launch_bits_from_pairs(const int32_t *ht, int32_t len, uint32_t *bits, uint16_t *vals, int32_t vmin,
                            uint32_t *partials, uint32_t *bad, hipStream_t s);

// --------------------------------------------------------- allocation map

// Every device allocation the engine hands to kernels (table columns, memory
// manager regions) is registered here; each launch's buffers are checked
// against it on the host before the launch, so a wrong pointer or a too
// short buffer becomes an error instead of a GPU memory fault.
class AllocRegistry {
  public:
    void add(const void *p, uint64_t bytes);
    void remove(const void *p);
    // throws unless [p, p + bytes) lies inside one registered allocation
    void check(const void *p, uint64_t bytes, const char *what) const;

  private:
    std::map<uintptr_t, uint64_t> allocs_;
};

// ----------------------------------------------------------- memory manager

// Bump allocator over device regions (operations/memory_manager.hpp:61-164).
// alloc_zero is served from regions that are zero when handed out; reset()
// rewinds every region and re-zeroes the used part of the zero regions
// (memory_manager.cpp:42-53, 71-90).  Throws std::bad_alloc past the budget.
class MemoryManager {
  public:
    MemoryManager(uint64_t budget, hipStream_t stream, AllocRegistry *registry, uint64_t max_region = 1ull << 30);
    ~MemoryManager();
    MemoryManager(const MemoryManager &) = delete;
    MemoryManager &operator=(const MemoryManager &) = delete;

    template <class T> T *alloc(uint64_t count) { return static_cast<T *>(raw(count * sizeof(T), false)); }
    template <class T> T *alloc_zero(uint64_t count) { return static_cast<T *>(raw(count * sizeof(T), true)); }
    void reset(hipStream_t s);
    uint64_t used_bytes() const;

  private:
    struct Region {
        uint8_t *base;
        uint64_t size, used;
    };
    void *raw(uint64_t bytes, bool zero);
    std::vector<Region> regions_, zero_regions_;
    hipStream_t stream_;
    AllocRegistry *registry_;
    uint64_t budget_, budget_zero_, reserved_ = 0, reserved_zero_ = 0, max_region_;
};

// ------------------------------------------------------------------ columns

// A column of int32 (or, for aggregate results, uint64) values split into
// segments of `segment_size` rows with per-segment min/max
// (models/models.hpp:20-1649).  Storage is one device buffer; segment s is
// the view starting at row s * segment_size.
struct Column {
    uint64_t nrows = 0;
    uint64_t segment_size = 1;
    bool loaded = false;        // false: a column the base table does not hold
    bool is_aggregate = false;  // values are uint64
    void *data = nullptr;       // device storage
    std::vector<int32_t> seg_min, seg_max;

    uint64_t nseg() const { return nrows == 0 ? 0 : (nrows + segment_size - 1) / segment_size; }
    uint64_t seg_rows(uint64_t s) const { return std::min(segment_size, nrows - s * segment_size); }
    int32_t *i32(uint64_t seg = 0) const { return static_cast<int32_t *>(data) + seg * segment_size; }
    uint64_t *u64(uint64_t seg = 0) const { return static_cast<uint64_t *>(data) + seg * segment_size; }
    std::pair<int32_t, int32_t> get_min_max() const; // models.hpp:1257-1272
};

struct Table {
    std::string name;
    uint64_t nrows = 0;
    std::vector<Column> columns;            // table_column_numbers entries
    std::vector<void *> owned;              // device buffers to free
};

// ------------------------------------------------------------ kernel bundles

// KernelType, models/execution.hpp:30-46
enum class KernelType : uint8_t {
    EmptyKernel,
    LogicalKernel,
    SelectionKernelColumns,
    SelectionKernelLiteral,
    FillKernel,
    PerformOperationKernelColumns,
    PerformOperationKernelLiteralFirst,
    PerformOperationKernelLiteralSecond,
    BuildKeysHTKernel,
    FilterJoinKernel,
    BuildKeyValsHTKernel,
    FullJoinKernel,
    AggregateOperationKernel,
    GroupByAggregateKernel,
};

struct GroupBySpec {
    std::vector<const int32_t *> keys;   // per key column (segment view)
    std::vector<int32_t> mins, maxs;
    std::vector<int32_t *> out_keys;
    uint64_t *sums = nullptr;
    uint32_t *seen = nullptr;
    uint64_t prod = 0;
};

// Compact probe tables of one join, built on the device when the join is
// applied.  Their statistics (set slots, whether every value fits 16 bits)
// decide the value layout, the LDS dictionary and the predicate order; the
// host reads them when a fused program first needs them, one stream sync
// for every join of the plan so far (Engine::resolve_compact).
struct CompactState {
    // built
    const uint32_t *bits = nullptr;
    const uint32_t *stats = nullptr; // device: value overflow, then nparts per-block counts of set slots
    uint32_t nparts = 0;
    bool pairs = false;
    const void *ht = nullptr;       // reference-layout table
    const uint16_t *narrow = nullptr; // 16-bit values, when the value range fits
    uint8_t *dict = nullptr;        // rank dictionary buffer (capacity dict_cap bytes), or null
    uint32_t dict_cap = 0;
    uint64_t words = 0, build_rows = 0;
    // resolved
    bool resolved = false;
    float sel = 1.0f;
    const void *vals = nullptr;
    int vkind = 0;
    const void *rank = nullptr;
    uint32_t rank_bytes = 0;
};

// One functor of the reference (the constructor arguments of the classes in
// kernels/*.hpp), with device pointers already offset to its segment.
struct KernelData {
    KernelType type = KernelType::EmptyKernel;
    int64_t n = 0;
    int cmp = 0, logic = 0, op = 0;
    int32_t literal = 0;
    const uint8_t *in_flags = nullptr;
    uint8_t *out_flags = nullptr;
    const uint8_t *local_flags = nullptr;
    const int32_t *a = nullptr, *b = nullptr;
    int32_t *result = nullptr;
    void *ht = nullptr;
    int32_t ht_len = 0, ht_min = 0, ht_max = 0;
    uint64_t *agg_res = nullptr;
    std::shared_ptr<GroupBySpec> gb;
    // Compact probe tables derived from `ht` once the build side ran (used by
    // fused bundles only): presence bitmap over [ht_min, ht_max] and, for a
    // full join, the value per slot (uint16 offset from vmin, or the pairs).
    const uint32_t *bits = nullptr;
    const void *vals = nullptr;
    int vkind = 0;       // FValKind
    const void *rank = nullptr; // rank dictionary of bits + vals (FV_RANK16 layout), if small
    uint32_t rank_bytes = 0;
    int32_t vmin = 0;
    float sel = 1.0f;    // estimated fraction of rows a predicate keeps
    // probes: the compact tables' statistics, read on first use (they then
    // supply vals, vkind, rank and sel)
    std::shared_ptr<CompactState> compact;
};

struct Engine;

// Buffers still referenced after a bundle: their final contents must be
// written back by a fused execution.
struct LiveSet {
    std::vector<std::pair<uintptr_t, uintptr_t>> ranges; // [begin, end)
    bool contains(const void *p) const
    {
        const uintptr_t a = (uintptr_t)p;
        for (auto &r : ranges)
            if (a >= r.first && a < r.second) return true;
        return false;
    }
    void add(const void *p, uint64_t bytes)
    {
        if (p && bytes) ranges.emplace_back((uintptr_t)p, (uintptr_t)p + bytes);
    }
};

struct ExecContext {
    Engine *eng;
    hipStream_t stream;
    const LiveSet *live;
    const std::string *table = nullptr; // base table the bundle runs over
    mutable uint64_t segment = 0;       // segment the bundle covers
    // flags buffer the table has not written yet: logically all true
    const uint8_t *virgin_begin = nullptr, *virgin_end = nullptr;
    mutable bool virgin_filled = false; // set once it holds values (filled, or stored whole by a program)
    // the all-true buffer stays in use after the bundles: it must end up
    // holding values (stored whole by a program, else filled)
    bool virgin_live = false;
};

// KernelBundle, models/execution.hpp:500-651: the operators queued for one
// segment, run in order on one stream.  With fusion on, the bundle becomes a
// single fused kernel when its operators allow it.
struct CapturedBundle;

class KernelBundle {
  public:
    void add_kernel(const KernelData &k) { kernels_.push_back(k); }
    const std::vector<KernelData> &get_kernels() const { return kernels_; }
    void execute(const ExecContext &ctx) const;
    // Builds and validates the fused program of a captured fact-table bundle
    // (throws NotBatchable when it does not fuse or touches a row-sized
    // intermediate).  The engine's capture must describe the batch.
    static void build_captured(Engine &eng, CapturedBundle &cb);

  private:
    void execute_unfused(const ExecContext &ctx) const;
    bool build_program(const ExecContext &ctx, FProgram &p) const;
    std::vector<KernelData> kernels_;
};

// ------------------------------------------------------------------- engine

// A pipeline that cannot take part in a batched fact scan (it runs through
// the plain path instead).
struct NotBatchable : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Capture of one pipeline's fact-table bundles for a batched execution
// (sdb_engine_execute_batch): while active, the fused program of every
// bundle over `table` is recorded instead of launched, the table's row-sized
// intermediates get placeholder addresses (never registered, so a launch
// that would touch one fails its bounds check), and device work that must
// follow the fact pass is deferred.
struct CapturedBundle {
    FProgram p;
    int64_t n;
    uint64_t segment;
    // The program is built once every pipeline of the batch has been
    // captured (KernelBundle::build_captured), so that the join statistics
    // of all of them are read in one stream sync; until then, the bundle and
    // what its build reads from the execution context.
    std::vector<KernelData> kernels;
    LiveSet live;
    std::string table;
    const uint8_t *virgin_begin = nullptr, *virgin_end = nullptr;
    bool virgin_filled = false;
};

struct Capture {
    bool active = false;
    std::string table;
    uint32_t lds_budget = F_LDS_BUDGET;           // LDS one captured program may use
    std::vector<CapturedBundle> bundles;
    std::vector<std::function<void()>> deferred;  // run after the batched fact pass
    uint64_t phantom_next = 0;
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // rows per segment: the reference uses 2^28 (common.hpp:12); with 288 GB
    // of HBM a segment here defaults to 2^30 rows, so a fact table of up to
    // ~1.07e9 rows is one launch per bundle (results do not depend on it)
    uint64_t segment_size = 1ull << 30;
    bool fusion = true;
    int codegen = SDB_CODEGEN_AUTO; // how fused kernels are produced
    int unroll = 0;                 // chunks per thread per iteration (0: default)
    AllocRegistry registry;
    std::unique_ptr<MemoryManager> mm;
    // compact tables whose statistics the host has not read yet
    std::vector<std::weak_ptr<CompactState>> pending_compact;
    // Reads the statistics of every pending compact table (one stream sync).
    void resolve_compact();
    std::map<std::string, std::unique_ptr<Table>> tables;
    int64_t launches = 0;
    double last_device_ms = 0, last_wall_ms = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // pinned host staging of result downloads (grown on demand, freed with
    // the engine): one asynchronous copy per array and one sync
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    void *pinned_buffer(size_t bytes);
    // Waits until the stream's work so far is done, polling an event: the
    // engine waits only between device phases that follow each other
    // closely, where a sleeping wait adds the thread's wake-up latency
    // (measured up to ~0.5 ms per query step).
    hipEvent_t sync_ev = nullptr;
    void wait_stream();
    // While a plan executes, the waits record ev_stop instead of sync_ev: the
    // plan's last wait (its result transfer) then marks its end on the device
    // with no further event round trip (stop_recorded: one has been).
    bool stop_at_wait = false, stop_recorded = false;

    // Optional per-launch timing with HIP events on the engine stream.
    struct KernelStat {
        int64_t launches = 0, rows = 0;
        double ms = 0;
    };
    bool profile = false;
    std::string prof_table; // when set, only the bundles of this table are timed
    bool profiles(const std::string *table) const
    {
        return profile && (prof_table.empty() || (table && *table == prof_table));
    }
    std::map<std::string, KernelStat> prof_stats;
    struct ProfRec {
        std::string name;
        int64_t rows;
        hipEvent_t a, b;
    };
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> event_pool;
    void prof_begin(const std::string &name, int64_t rows, ProfRec &rec);
    void prof_end(ProfRec &rec);
    void prof_collect(); // after the stream is synchronised

    Table *find_table(const std::string &name);

    // SDB_TRACE_STEPS=1: host timestamps of an execution's phases, printed
    // to stderr when it ends (where host time goes between the kernels)
    bool tracing = false;
    std::vector<std::pair<std::chrono::steady_clock::time_point, std::string>> marks;
    void mark(const std::string &what)
    {
        if (tracing) marks.emplace_back(std::chrono::steady_clock::now(), what);
    }
    void print_marks();

    Capture capture;
    // Buffer of `count` T for a row-sized intermediate of `table` (a
    // placeholder address while that table's bundles are being captured).
    template <class T> T *row_buffer(const std::string &table, uint64_t count)
    {
        if (capture.active && table == capture.table) return static_cast<T *>(phantom(count * sizeof(T)));
        return mm->alloc<T>(count);
    }
    void *phantom(uint64_t bytes);
};

// -------------------------------------------------------- transient table

// TransientTable, models/transient_table.hpp:20-1928, for one device.
class TransientTable {
  public:
    TransientTable(Engine &eng, Table *base);

    void set_group_by_column(uint64_t col); // 189-193
    void apply_filter(const ExprType &expr, const std::string &parent_op);
    void apply_project(const std::vector<ExprType> &exprs);
    void apply_aggregate(const AggType &agg, const std::vector<int64_t> &group);
    void apply_join(TransientTable &right, const RelNode &rel);
    // Runs the queued bundles (383-506).  `dead_after`: the caller replaces
    // the table's columns and flags right after, so nothing needs writing
    // back.
    void execute_pending_kernels(bool dead_after = false);

    uint64_t get_nrows() const { return nrows_; }
    const std::vector<Column *> &get_columns() const { return current_columns_; }
    // Device flags of the table (materialised if still all-true).
    const uint8_t *flags_device();

  private:
    // throws unless the flags buffer is registered device memory (a captured
    // fact table's placeholder is not: NotBatchable)
    void check_flags_memory() const;
    uint64_t segment_num() const { return nrows_ == 0 ? 0 : (nrows_ + eng_.segment_size - 1) / eng_.segment_size; }
    uint64_t seg_rows(uint64_t s) const { return std::min(eng_.segment_size, nrows_ - s * eng_.segment_size); }
    void add_pending_kernels(std::vector<KernelBundle> ops); // 347-381
    void ensure_secondary_flags();
    void swap_flag_buffers();
    Column &new_column(uint64_t nrows, bool aggregate = false);
    Column *col(int64_t idx, const char *what) const;
    std::tuple<uint8_t *, int32_t, int32_t> build_keys_hash_table(int column);     // 707-726
    std::tuple<int32_t *, int32_t, int32_t> build_key_vals_hash_table(int column); // 728-748
    void apply_search(const ExprType &expr, const std::string &parent_op);

    Engine &eng_;
    uint64_t nrows_;
    uint8_t *flags_ = nullptr, *flags_secondary_ = nullptr;
    uint8_t *virgin_ = nullptr;          // flags buffer that is logically all
                                         // true and not yet written
    // estimated fraction of rows the filters so far keep (from the column
    // statistics; it orders fused predicates, never decides a result)
    float est_kept_ = 1.0f;
    void note_filter(int logic, float sel);
    std::vector<Column *> current_columns_;
    std::deque<Column> materialized_columns_;
    Column *group_by_column_ = nullptr;
    uint64_t group_by_column_index_ = 0;
    std::vector<std::vector<KernelBundle>> pending_kernels_; // [phase][segment]
    std::string name_;
};

// Host-side bounds checks of one launch (validate.cpp).
void validate_kernel(const KernelData &k, const AllocRegistry &reg);
void validate_program(const FProgram &p, int64_t n, const AllocRegistry &reg);

// ------------------------------------------------------------------ executor

struct ExecutionInfo {
    std::map<std::string, std::vector<int>> loaded_columns;
    std::map<std::string, int> group_by_columns;
    std::vector<int> dag_order;
};

std::vector<int> dag_topological_sort(const PlanResult &plan);   // preprocessing.cpp:14-102
ExecutionInfo parse_execution_info(const PlanResult &plan);     // preprocessing.cpp:104-275

// Host copy of a final table.
// Host copy of a final table.  Dense form: flags and every column hold nrows
// entries.  Sparse form (large group-by tables, compacted on the device):
// `sel` holds the selected rows in ascending order and every column only its
// values at those rows; flags are empty and unselected rows read as 0 (the
// sums of unselected group-by slots are 0, their keys are never written).
struct ResultTable {
    uint64_t nrows = 0;
    bool sparse = false;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> sel;
    struct Col {
        bool is_u64;
        std::vector<int32_t> i32;
        std::vector<uint64_t> u64;
    };
    std::vector<Col> cols;

    uint64_t selected() const;
    void flags_to(uint8_t *host) const;                 // nrows flags
    void col_to(int c, void *host) const;               // nrows values
    void selected_rows_to(uint64_t *host) const;        // selected() row ids
    void col_selected_to(int c, void *host) const;      // selected() values
};

// The final table of a plan, on the device (valid until the allocator resets).
struct FinalTable {
    uint64_t nrows = 0;
    const uint8_t *flags = nullptr;
    struct Col {
        const void *data;
        bool is_u64;
        uint64_t rows;
    };
    std::vector<Col> cols;
};

// execute_ddor_plan, executor/ddor_executor.cpp:22-245, in two steps: the
// plan walk (device work queued on the engine stream) and the host copy of
// the final table (result_writer.cpp's materialize_host_view).
FinalTable run_ddor_plan(Engine &eng, const PlanResult &plan);
ResultTable copy_result(Engine &eng, const FinalTable &t);
std::vector<ResultTable> copy_results(Engine &eng, const std::vector<FinalTable> &ts);
ResultTable execute_ddor_plan(Engine &eng, const PlanResult &plan);
// Several plans with a shared fact scan (sdb_engine_execute_batch); the
// results are in plan order.  *batched = plans that shared the scan.
std::vector<ResultTable> execute_batch(Engine &eng, const std::vector<PlanResult> &plans, int *batched);

extern const std::map<std::string, int> TABLE_COLUMN_NUMBERS;   // kernels/types.cpp:13-21
extern const std::map<std::string, std::vector<int>> TABLE_COLUMN_INDICES; // types.cpp:3-11

} // namespace sdb
