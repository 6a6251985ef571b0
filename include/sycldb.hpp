// sycldb.hpp -- C++ layer over the C ABI (sycldb.h) with the reference's own
// operator interface: the free functions of kernels/selection.hpp,
// kernels/join.hpp, kernels/aggregation.hpp and kernels/common.hpp under the
// same names, argument order and meaning, and the same error behaviour
// (std::invalid_argument for an unknown arithmetic op, std::bad_alloc when
// memory runs out).  A reference translation unit switches by including this
// header instead of those four and passing a sycldb::Queue where it passed a
// sycl::queue.
//
// Differences, by design:
//   - sycl::queue / sycl::event become sycldb::Queue (a HIP stream) and
//     sycldb::Event.  Work on one queue runs in submission order, so
//     dependencies on events of the same queue are met by that order; an
//     event of another queue is waited for before the submission.
//   - bool arrays are passed through as 1-byte flags (sizeof(bool) == 1, as
//     the reference assumes).
//   - group_by_aggregate's pointer tables (contents, results) and its
//     min / max arrays must be readable by the host: the reference allocates
//     them with malloc_shared.
// Header-only; needs only the C ABI (link libsycldb_hip.so).
#ifndef SYCLDB_HPP
#define SYCLDB_HPP

#include <cctype>
#include <cstdint>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "sycldb.h"

namespace sycldb {

typedef sdb_comp_op comp_op;       // kernels/selection.hpp:7-15
typedef sdb_logical_op logical_op; // kernels/selection.hpp:16-21
typedef sdb_binop BinaryOp;        // kernels/aggregation.hpp:11-17

// Status of a C-ABI call -> the reference's exception types.
inline void check(int status)
{
    if (status == SDB_OK) return;
    const std::string msg = sdb_last_error();
    if (status == SDB_ERR_INVALID_ARG) throw std::invalid_argument(msg);
    if (status == SDB_ERR_OOM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

class Queue {
  public:
    explicit Queue(sdb_stream stream = nullptr) : stream_(stream) {}
    sdb_stream stream() const { return stream_; }
    void wait() const { check(sdb_stream_synchronize(stream_)); }

  private:
    sdb_stream stream_;
};

// Completion of the work submitted to a queue up to this point.
class Event {
  public:
    Event() = default;
    explicit Event(const Queue &q) : stream_(q.stream()), valid_(true) {}
    void wait() const
    {
        if (valid_) check(sdb_stream_synchronize(stream_));
    }
    void wait_and_throw() const { wait(); }
    sdb_stream stream() const { return stream_; }

  private:
    sdb_stream stream_ = nullptr;
    bool valid_ = false;
};

namespace detail {
inline void depend(const Queue &q, const std::vector<Event> &deps)
{
    for (const Event &e : deps)
        if (e.stream() != q.stream()) e.wait();
}
inline const uint8_t *u8(const bool *p) { return reinterpret_cast<const uint8_t *>(p); }
inline uint8_t *u8(bool *p) { return reinterpret_cast<uint8_t *>(p); }
} // namespace detail

// ----------------------------------------------------- kernels/types.hpp

// ColumnData / TableData, kernels/types.hpp:12-37 (same fields): the
// operand descriptions of the TableData / ColumnData overloads of
// full_join and group_by_aggregate.  content and flags are device pointers.
template <typename T> struct ColumnData {
    T *content;
    bool has_ownership;
    bool is_aggregate_result;
    T min_value;
    T max_value;
};

template <typename T> struct TableData {
    ColumnData<T> *columns;
    int columns_size;
    int col_len;
    int col_number;
    bool *flags;
    int group_by_column;
    void *ht;
    int ht_min, ht_max;
    std::string table_name;
    std::map<int, int> column_indices;
};

// A device allocator with the memory_manager interface the allocating
// overloads use (operations/memory_manager.hpp:79-82): alloc<T>(count,
// on_device) and alloc_zero<T>(count).  Frees everything it handed out when
// destroyed; std::bad_alloc when the device is out of memory.
class DeviceAllocator {
  public:
    DeviceAllocator() = default;
    DeviceAllocator(const DeviceAllocator &) = delete;
    DeviceAllocator &operator=(const DeviceAllocator &) = delete;
    ~DeviceAllocator()
    {
        for (void *p : held_) sdb_device_free(p);
    }
    template <class T> T *alloc(uint64_t count, bool on_device = true)
    {
        (void)on_device;
        void *p = nullptr;
        check(sdb_device_alloc(&p, count * sizeof(T) ? count * sizeof(T) : 1));
        held_.push_back(p);
        return static_cast<T *>(p);
    }
    template <class T> T *alloc_zero(uint64_t count)
    {
        // rounded up to whole int32 words, cleared by the fill kernel
        const uint64_t words = (count * sizeof(T) + 3) / 4;
        int32_t *p = alloc<int32_t>(words ? words : 1);
        check(sdb_fill_i32(p, 0, (int64_t)words, nullptr));
        check(sdb_stream_synchronize(nullptr));
        return reinterpret_cast<T *>(p);
    }

  private:
    std::vector<void *> held_;
};

// ------------------------------------------------------- fused bundles

// KernelBundle::execute with its fused path on (models/execution.hpp:
// 528-650): the operators, over n rows, as one kernel when they fuse.
inline Event bundle_execute(const std::vector<sdb_op> &ops, int64_t n, Queue &queue,
                            const std::vector<Event> &dependencies = {})
{
    detail::depend(queue, dependencies);
    check(sdb_bundle_exec(ops.data(), (int)ops.size(), n, queue.stream()));
    return Event(queue);
}

// ------------------------------------------------- kernels/selection.hpp

inline comp_op get_comp_op(std::string op) { return (comp_op)sdb_comp_op_from_string(op.c_str()); }

inline logical_op get_logical_op(std::string op) { return (logical_op)sdb_logical_op_from_string(op.c_str()); }

// any op made of letters is a logical one (kernels/selection.hpp:53-59)
inline bool is_filter_logical(const std::string &op)
{
    for (char c : op)
        if (!std::isalpha((unsigned char)c)) return false;
    return true;
}

// flags[i] = logical(parent_op, flags[i], compare(op, arr[i], value))
inline Event selection(bool flags[], const int arr[], std::string op, int value, std::string parent_op, int col_len,
                       Queue &queue, const std::vector<Event> deps)
{
    detail::depend(queue, deps);
    check(sdb_select_lit(arr, detail::u8(flags), detail::u8(flags), col_len, get_comp_op(op),
                         get_logical_op(parent_op), value, queue.stream()));
    return Event(queue);
}

// flags[i] = logical(parent_op, flags[i], compare(op, operand1[i], operand2[i]))
inline Event selection(bool flags[], const int operand1[], std::string op, const int operand2[],
                       std::string parent_op, int col_len, Queue &queue, const std::vector<Event> deps)
{
    detail::depend(queue, deps);
    check(sdb_select_col(operand1, operand2, detail::u8(flags), detail::u8(flags), col_len, get_comp_op(op),
                         get_logical_op(parent_op), queue.stream()));
    return Event(queue);
}

// ------------------------------------------------------ kernels/join.hpp

inline Event build_keys_ht(const int col[], const bool flags[], int col_len, bool ht[], int ht_len, int ht_min_value,
                           Queue &queue, const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_build_keys(col, detail::u8(flags), col_len, detail::u8(ht), ht_len, ht_min_value, queue.stream()));
    return Event(queue);
}

inline Event build_key_vals_ht(int col[], int agg_col[], const bool flags[], int col_len, int ht[], int ht_len,
                               int ht_min_value, Queue &queue, const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_build_keyvals(col, agg_col, detail::u8(flags), col_len, ht, ht_len, ht_min_value, queue.stream()));
    return Event(queue);
}

// probe_col_flags[i] = flag && min <= probe <= max && build_ht[probe - min]
inline Event filter_join(const int *probe_col, bool *probe_col_flags, int probe_col_len, const bool *build_ht,
                         int build_min_value, int build_max_value, Queue &queue,
                         const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_probe_semi(probe_col, detail::u8(probe_col_flags), detail::u8(probe_col_flags), probe_col_len,
                         detail::u8(build_ht), build_min_value, build_max_value, queue.stream()));
    return Event(queue);
}

// Build (when build_ht is null: a zeroed table from the allocator, filled
// by build_keys_ht) and probe, kernels/join.hpp:210-254.
template <class Allocator>
inline Event filter_join(int build_col[], bool build_flags[], int build_col_len, int build_max_value,
                         int build_min_value, int probe_col[], bool probe_col_flags[], int probe_col_len,
                         bool *build_ht, Allocator &gpu_allocator, Queue &queue,
                         const std::vector<Event> &dependencies)
{
    std::vector<Event> events(dependencies);
    const int ht_len = build_max_value - build_min_value + 1;
    bool *ht = build_ht;
    if (ht == nullptr) {
        ht = gpu_allocator.template alloc_zero<bool>(ht_len);
        events = {build_keys_ht(build_col, build_flags, build_col_len, ht, ht_len, build_min_value, queue, events)};
    }
    return filter_join(probe_col, probe_col_flags, probe_col_len, ht, build_min_value, build_max_value, queue, events);
}

inline Event full_join(const int *probe_col, int *probe_val_out, bool *probe_flags, int probe_col_len, const int *ht,
                       int ht_min, int ht_max, Queue &queue, const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_probe_full(probe_col, probe_val_out, detail::u8(probe_flags), detail::u8(probe_flags), probe_col_len,
                         ht, ht_min, ht_max, queue.stream()));
    return Event(queue);
}

// The TableData overload, kernels/join.hpp:357-461: builds the build
// table's key-value table (unless build_table.ht holds one) over its join
// column's [min, max] with its group-by column as the value, probes the
// probe table's join column in place (the joined value replaces the key),
// then renames that column to the build group-by column's position and gives
// it that column's min / max.
template <class Allocator>
inline Event full_join(TableData<int> &probe_table, TableData<int> &build_table, int probe_col_index,
                       int build_col_index, Allocator &gpu_allocator, Queue &queue,
                       const std::vector<Event> &dependencies)
{
    std::vector<Event> events(dependencies);
    const int build_column = build_table.column_indices.at(build_col_index),
              probe_column = probe_table.column_indices.at(probe_col_index),
              group_by_column = build_table.column_indices.at(build_table.group_by_column);
    int build_col_min, build_col_max, *ht;
    if (build_table.ht != nullptr) {
        ht = static_cast<int *>(build_table.ht);
        build_col_min = build_table.ht_min;
        build_col_max = build_table.ht_max;
    } else {
        build_col_min = build_table.columns[build_column].min_value;
        build_col_max = build_table.columns[build_column].max_value;
        const int ht_len = build_col_max - build_col_min + 1;
        ht = gpu_allocator.template alloc_zero<int>((uint64_t)ht_len * 2);
        events = {build_key_vals_ht(build_table.columns[build_column].content,
                                    build_table.columns[group_by_column].content, build_table.flags,
                                    build_table.col_len, ht, ht_len, build_col_min, queue, events)};
    }
    int *probe_content = probe_table.columns[probe_column].content;
    Event e = full_join(probe_content, probe_content, probe_table.flags, probe_table.col_len, ht, build_col_min,
                        build_col_max, queue, events);
    probe_table.column_indices.erase(probe_col_index);
    probe_table.column_indices[probe_table.col_number + build_table.group_by_column] = probe_column;
    probe_table.columns[probe_column].min_value = build_table.columns[group_by_column].min_value;
    probe_table.columns[probe_column].max_value = build_table.columns[group_by_column].max_value;
    return e;
}

// ----------------------------------------------- kernels/aggregation.hpp

inline BinaryOp get_op_from_string(const std::string &op)
{
    const int v = sdb_binop_from_string(op.c_str());
    if (v < 0) throw std::invalid_argument("Unsupported operation: " + op);
    return (BinaryOp)v;
}

// result[i] = a[i] op b[i] where flags[i]
inline Event perform_operation(int result[], const int a[], const int b[], const bool flags[], int size,
                               const std::string &op, Queue &queue, const std::vector<Event> &dependencies)
{
    const BinaryOp o = get_op_from_string(op);
    detail::depend(queue, dependencies);
    check(sdb_binop_cc(result, a, b, detail::u8(flags), size, o, queue.stream()));
    return Event(queue);
}

inline Event perform_operation(int result[], int a, const int b[], const bool flags[], int size, const std::string &op,
                               Queue &queue, const std::vector<Event> &dependencies)
{
    const BinaryOp o = get_op_from_string(op);
    detail::depend(queue, dependencies);
    check(sdb_binop_lc(result, a, b, detail::u8(flags), size, o, queue.stream()));
    return Event(queue);
}

inline Event perform_operation(int result[], const int a[], int b, const bool flags[], int size, const std::string &op,
                               Queue &queue, const std::vector<Event> &dependencies)
{
    const BinaryOp o = get_op_from_string(op);
    detail::depend(queue, dependencies);
    check(sdb_binop_cl(result, a, b, detail::u8(flags), size, o, queue.stream()));
    return Event(queue);
}

// *agg_res += sum of (uint64)(int64)a[i] over flagged rows (no reset)
inline Event aggregate_operation(const int a[], const bool flags[], int size, uint64_t *agg_res, Queue &queue,
                                 const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_sum_flagged(a, detail::u8(flags), size, agg_res, queue.stream()));
    return Event(queue);
}

// Dense group-by SUM into agg_result / result_flags / results (all zeroed
// by the caller's allocator), kernels/aggregation.hpp:346-409.
inline Event group_by_aggregate(const int **contents, const int *agg_column, const int *max, const int *min,
                                const bool *flags, int col_len, int col_num, int **results, uint64_t *agg_result,
                                unsigned *result_flags, uint64_t prod_ranges, const std::string &agg_op,
                                Queue &gpu_queue, const std::vector<Event> &dependencies)
{
    if (agg_op != "SUM") throw std::invalid_argument("Unsupported aggregate operation: " + agg_op);
    detail::depend(gpu_queue, dependencies);
    check(sdb_groupby_dense(contents, col_num, min, max, agg_column, detail::u8(flags), col_len, results, agg_result,
                            result_flags, prod_ranges, gpu_queue.stream()));
    return Event(gpu_queue);
}

// The ColumnData overload, kernels/aggregation.hpp:411-539: allocates the
// key columns (results), sums and seen flags of prod = product of the key
// ranges slots from the allocator, aggregates, and turns the seen flags into
// row flags.  Returns (results, prod, flags, sums, event).  `results` is a
// host array of col_num device pointers from new[] (the reference uses
// malloc_shared): the caller releases it with delete[]; the device arrays
// belong to the allocator.
template <class Allocator>
inline std::tuple<int **, unsigned long long, bool *, uint64_t *, Event>
group_by_aggregate(ColumnData<int> *group_columns, int *agg_column, bool *flags, int col_num, int col_len,
                   const std::string &agg_op, Allocator &gpu_allocator, Queue &queue,
                   const std::vector<Event> &dependencies)
{
    if (agg_op != "SUM") throw std::invalid_argument("Unsupported aggregate operation: " + agg_op);
    unsigned long long prod_ranges = 1;
    for (int i = 0; i < col_num; i++)
        prod_ranges *= group_columns[i].max_value - group_columns[i].min_value + 1;
    std::unique_ptr<int *[]> results(new int *[col_num]);
    std::vector<const int *> contents(col_num);
    std::vector<int> mins(col_num), maxs(col_num);
    for (int i = 0; i < col_num; i++) {
        results[i] = gpu_allocator.template alloc<int>(prod_ranges, true);
        contents[i] = group_columns[i].content;
        mins[i] = group_columns[i].min_value;
        maxs[i] = group_columns[i].max_value;
    }
    uint64_t *agg_result = gpu_allocator.template alloc_zero<uint64_t>(prod_ranges);
    unsigned *res_flags = gpu_allocator.template alloc_zero<unsigned>(prod_ranges);
    bool *final_flags = gpu_allocator.template alloc<bool>(prod_ranges, true);
    detail::depend(queue, dependencies);
    check(sdb_groupby_dense(contents.data(), col_num, mins.data(), maxs.data(), agg_column, detail::u8(flags), col_len,
                            results.get(), agg_result, res_flags, prod_ranges, queue.stream()));
    check(sdb_seen_to_flags(res_flags, detail::u8(final_flags), prod_ranges, queue.stream()));
    return std::make_tuple(results.release(), prod_ranges, final_flags, agg_result, Event(queue));
}

// ---------------------------------------------------- kernels/common.hpp

// count_true_flags_async, kernels/common.hpp:18-65: the count lands in
// device memory owned by the result; get() waits and reads it.
struct AsyncCountResult {
    std::shared_ptr<uint64_t> result;
    Event event;
    uint64_t get()
    {
        event.wait_and_throw();
        uint64_t h = 0;
        check(sdb_copy_to_host(&h, result.get(), sizeof(h), event.stream()));
        return h;
    }
};

inline AsyncCountResult count_true_flags_async(const bool *flags, int len, Queue &queue,
                                               const std::vector<Event> &dependencies = {})
{
    detail::depend(queue, dependencies);
    uint64_t *d = nullptr;
    check(sdb_device_alloc(reinterpret_cast<void **>(&d), sizeof(uint64_t)));
    std::shared_ptr<uint64_t> count(d, [](uint64_t *p) {
        if (p) sdb_device_free(p);
    });
    check(sdb_count_flags(detail::u8(flags), len, d, 0, queue.stream()));
    return {count, Event(queue)};
}

// number of set flags (waits for the result), kernels/common.hpp:67-79
inline uint64_t count_true_flags(const bool *flags, int len, Queue &queue, const std::vector<Event> &dependencies = {})
{
    return count_true_flags_async(flags, len, queue, dependencies).get();
}

// *result += number of set flags (the allocator overload, common.hpp:81-110;
// the allocator argument is accepted for the signature and not used)
template <class Allocator>
inline Event count_true_flags(const bool *flags, int len, Queue &queue, Allocator &allocator, uint64_t *result,
                              const std::vector<Event> &dependencies = {})
{
    (void)allocator;
    detail::depend(queue, dependencies);
    check(sdb_count_flags(detail::u8(flags), len, result, 1, queue.stream()));
    return Event(queue);
}

} // namespace sycldb

#endif
