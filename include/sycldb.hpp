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
#include <new>
#include <stdexcept>
#include <string>
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

inline Event full_join(const int *probe_col, int *probe_val_out, bool *probe_flags, int probe_col_len, const int *ht,
                       int ht_min, int ht_max, Queue &queue, const std::vector<Event> &dependencies)
{
    detail::depend(queue, dependencies);
    check(sdb_probe_full(probe_col, probe_val_out, detail::u8(probe_flags), detail::u8(probe_flags), probe_col_len,
                         ht, ht_min, ht_max, queue.stream()));
    return Event(queue);
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

// ---------------------------------------------------- kernels/common.hpp

// number of set flags (waits for the result), kernels/common.hpp:67-79
inline uint64_t count_true_flags(const bool *flags, int len, Queue &queue, const std::vector<Event> &dependencies = {})
{
    detail::depend(queue, dependencies);
    uint64_t *d = nullptr;
    check(sdb_device_alloc(reinterpret_cast<void **>(&d), sizeof(uint64_t)));
    uint64_t h = 0;
    int st = sdb_count_flags(detail::u8(flags), len, d, 0, queue.stream());
    if (st == SDB_OK) st = sdb_copy_to_host(&h, d, sizeof(h), queue.stream());
    sdb_device_free(d);
    check(st);
    return h;
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
