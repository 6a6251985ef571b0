"""Operator-level API: the free functions of the reference's kernels/*.hpp on
torch device tensors, through the C ABI.

Names, argument order and meaning follow the reference:
  selection (kernels/selection.hpp:188-242), perform_operation
  (kernels/aggregation.hpp:74-206), build_keys_ht / build_key_vals_ht /
  filter_join / full_join (kernels/join.hpp:35-355), aggregate_operation and
  group_by_aggregate (kernels/aggregation.hpp:230-409), count_true_flags
  (kernels/common.hpp:30-110).
Flags are torch.uint8 tensors (0/1), columns torch.int32, sums torch.int64
holding the uint64 bit pattern.  Work is queued on torch's current stream.
Unknown arithmetic operators raise ValueError (the reference throws
std::invalid_argument); unknown comparison / logical operators map to EQ /
NONE as in the reference.
"""
import ctypes

import torch

from . import _lib as L

EQ, NE, LT, LE, GT, GE = L.EQ, L.NE, L.LT, L.LE, L.GT, L.GE
NONE, AND, OR = L.NONE, L.AND, L.OR


def get_comp_op(op):
    return L.lib.sdb_comp_op_from_string(op.encode())


def get_logical_op(op):
    return L.lib.sdb_logical_op_from_string(op.encode())


def get_op_from_string(op):
    v = L.lib.sdb_binop_from_string(op.encode())
    if v < 0:
        raise ValueError("Unknown operation: " + op)
    return v


def _stream(stream=None):
    if stream is not None:
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, dtype, name, n=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.device.type != "cuda":
        raise ValueError("%s must be a device tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if n is not None and t.numel() < n:
        raise ValueError("%s has %d elements, need %d" % (name, t.numel(), n))
    return ctypes.c_void_p(t.data_ptr())


def _n(n, *tensors):
    if n is None:
        n = tensors[0].numel()
    return int(n)


# ------------------------------------------------------------------ functors

def select_lit(col, in_flags, out_flags, cmp, logic, value, n=None, stream=None):
    """SelectionKernelLiteral (kernels/selection.hpp:157-186)."""
    n = _n(n, col)
    L.check(L.lib.sdb_select_lit(_dev(col, torch.int32, "col", n), _dev(in_flags, torch.uint8, "in", n),
                                 _dev(out_flags, torch.uint8, "out", n), n, cmp, logic, int(value),
                                 _stream(stream)))


def select_col(a, b, in_flags, out_flags, cmp, logic, n=None, stream=None):
    """SelectionKernelColumns (kernels/selection.hpp:127-155)."""
    n = _n(n, a)
    L.check(L.lib.sdb_select_col(_dev(a, torch.int32, "a", n), _dev(b, torch.int32, "b", n),
                                 _dev(in_flags, torch.uint8, "in", n), _dev(out_flags, torch.uint8, "out", n),
                                 n, cmp, logic, _stream(stream)))


def logical(in_flags, local_flags, out_flags, logic, n=None, stream=None):
    """LogicalKernel (kernels/selection.hpp:106-125)."""
    n = _n(n, in_flags)
    L.check(L.lib.sdb_logical(_dev(in_flags, torch.uint8, "in", n), _dev(local_flags, torch.uint8, "local", n),
                              _dev(out_flags, torch.uint8, "out", n), n, logic, _stream(stream)))


def fill(target, value, n=None, stream=None):
    """FillKernel (kernels/projection.hpp:7-21)."""
    n = _n(n, target)
    L.check(L.lib.sdb_fill_i32(_dev(target, torch.int32, "target", n), int(value), n, _stream(stream)))


def binop(result, a, b, flags, op, n=None, stream=None):
    """PerformOperationKernel{Columns,LiteralFirst,LiteralSecond}: a and b are
    int32 tensors or Python ints (at most one literal)."""
    if isinstance(a, int) and isinstance(b, int):
        raise ValueError("perform_operation needs at least one column")
    if isinstance(op, str):
        op = get_op_from_string(op)
    ref = b if isinstance(a, int) else a
    n = _n(n, ref)
    r = _dev(result, torch.int32, "result", n)
    f = _dev(flags, torch.uint8, "flags", n)
    s = _stream(stream)
    if isinstance(a, int):
        L.check(L.lib.sdb_binop_lc(r, a, _dev(b, torch.int32, "b", n), f, n, op, s))
    elif isinstance(b, int):
        L.check(L.lib.sdb_binop_cl(r, _dev(a, torch.int32, "a", n), b, f, n, op, s))
    else:
        L.check(L.lib.sdb_binop_cc(r, _dev(a, torch.int32, "a", n), _dev(b, torch.int32, "b", n), f, n, op, s))


def build_keys(keys, flags, ht, ht_len, ht_min, n=None, stream=None):
    n = _n(n, keys)
    L.check(L.lib.sdb_build_keys(_dev(keys, torch.int32, "keys", n), _dev(flags, torch.uint8, "flags", n), n,
                                 _dev(ht, torch.uint8, "ht", ht_len), int(ht_len), int(ht_min), _stream(stream)))


def build_keyvals(keys, vals, flags, ht, ht_len, ht_min, n=None, stream=None):
    n = _n(n, keys)
    L.check(L.lib.sdb_build_keyvals(_dev(keys, torch.int32, "keys", n), _dev(vals, torch.int32, "vals", n),
                                    _dev(flags, torch.uint8, "flags", n), n, _dev(ht, torch.int32, "ht", 2 * ht_len),
                                    int(ht_len), int(ht_min), _stream(stream)))


def probe_semi(probe, in_flags, out_flags, ht, ht_min, ht_max, n=None, stream=None):
    n = _n(n, probe)
    L.check(L.lib.sdb_probe_semi(_dev(probe, torch.int32, "probe", n), _dev(in_flags, torch.uint8, "in", n),
                                 _dev(out_flags, torch.uint8, "out", n), n,
                                 _dev(ht, torch.uint8, "ht", int(ht_max) - int(ht_min) + 1),
                                 int(ht_min), int(ht_max), _stream(stream)))


def probe_full(probe, val_out, in_flags, out_flags, ht, ht_min, ht_max, n=None, stream=None):
    n = _n(n, probe)
    L.check(L.lib.sdb_probe_full(_dev(probe, torch.int32, "probe", n), _dev(val_out, torch.int32, "val_out", n),
                                 _dev(in_flags, torch.uint8, "in", n), _dev(out_flags, torch.uint8, "out", n), n,
                                 _dev(ht, torch.int32, "ht", 2 * (int(ht_max) - int(ht_min) + 1)),
                                 int(ht_min), int(ht_max), _stream(stream)))


def sum_flagged(col, flags, acc, n=None, stream=None):
    """acc: int64 tensor of one element (uint64 bit pattern), accumulated into."""
    n = _n(n, col)
    L.check(L.lib.sdb_sum_flagged(_dev(col, torch.int32, "col", n), _dev(flags, torch.uint8, "flags", n), n,
                                  _dev(acc, torch.int64, "acc", 1), _stream(stream)))


def groupby_dense(keys, mins, maxs, agg, flags, out_keys, sums, seen, prod, n=None, stream=None):
    k = len(keys)
    n = _n(n, agg)
    key_ptrs = (ctypes.c_void_p * k)(*[_dev(c, torch.int32, "key", n).value for c in keys])
    out_ptrs = (ctypes.c_void_p * k)(*[_dev(c, torch.int32, "out_key", prod).value for c in out_keys])
    mins_a = (ctypes.c_int32 * k)(*mins)
    maxs_a = (ctypes.c_int32 * k)(*maxs)
    L.check(L.lib.sdb_groupby_dense(key_ptrs, k, mins_a, maxs_a, _dev(agg, torch.int32, "agg", n),
                                    _dev(flags, torch.uint8, "flags", n), n, out_ptrs,
                                    _dev(sums, torch.int64, "sums", prod), _dev(seen, torch.int32, "seen", prod),
                                    int(prod), _stream(stream)))


def seen_to_flags(seen, flags, prod, stream=None):
    L.check(L.lib.sdb_seen_to_flags(_dev(seen, torch.int32, "seen", prod), _dev(flags, torch.uint8, "flags", prod),
                                    int(prod), _stream(stream)))


def count_flags(flags, out, accumulate=False, n=None, stream=None):
    n = _n(n, flags)
    L.check(L.lib.sdb_count_flags(_dev(flags, torch.uint8, "flags", n), n, _dev(out, torch.int64, "out", 1),
                                  1 if accumulate else 0, _stream(stream)))


def compact(flags, row_ids, count, n=None, stream=None):
    n = _n(n, flags)
    L.check(L.lib.sdb_compact(_dev(flags, torch.uint8, "flags", n), n, _dev(row_ids, torch.int32, "row_ids", n),
                              _dev(count, torch.int32, "count", 1), _stream(stream)))


def gather(src, row_ids, count, dst, stream=None):
    L.check(L.lib.sdb_gather_i32(_dev(src, torch.int32, "src"), _dev(row_ids, torch.int32, "row_ids", count),
                                 int(count), _dev(dst, torch.int32, "dst", count), _stream(stream)))


def minmax(col, n=None, stream=None):
    n = _n(n, col)
    mn, mx = ctypes.c_int32(), ctypes.c_int32()
    L.check(L.lib.sdb_minmax_i32(_dev(col, torch.int32, "col", n), n, ctypes.byref(mn), ctypes.byref(mx),
                                 _stream(stream)))
    return mn.value, mx.value


# ------------------------------------------------- reference free functions

def selection(flags, arr, op, value, parent_op, col_len, stream=None):
    """selection(flags, arr, op, value, parent_op, col_len, queue, deps),
    kernels/selection.hpp:188-214 (in place on flags); `value` may also be a
    column (the overload at 216-242)."""
    cmp, logic = get_comp_op(op), get_logical_op(parent_op)
    if isinstance(value, torch.Tensor):
        select_col(arr, value, flags, flags, cmp, logic, n=col_len, stream=stream)
    else:
        select_lit(arr, flags, flags, cmp, logic, value, n=col_len, stream=stream)


def perform_operation(result, a, b, flags, size, op, stream=None):
    """perform_operation, kernels/aggregation.hpp:74-206."""
    binop(result, a, b, flags, get_op_from_string(op), n=size, stream=stream)


def build_keys_ht(col, flags, col_len, ht, ht_len, ht_min_value, stream=None):
    """build_keys_ht, kernels/join.hpp:35-58."""
    build_keys(col, flags, ht, ht_len, ht_min_value, n=col_len, stream=stream)


def build_key_vals_ht(col, agg_col, flags, col_len, ht, ht_len, ht_min_value, stream=None):
    """build_key_vals_ht, kernels/join.hpp:95-127."""
    build_keyvals(col, agg_col, flags, ht, ht_len, ht_min_value, n=col_len, stream=stream)


def filter_join(probe_col, probe_col_flags, probe_col_len, build_ht, build_min_value, build_max_value,
                stream=None):
    """filter_join with a pre-built table, kernels/join.hpp:175-208 (in place)."""
    probe_semi(probe_col, probe_col_flags, probe_col_flags, build_ht, build_min_value, build_max_value,
               n=probe_col_len, stream=stream)


def full_join(probe_col, probe_val_out, probe_flags, probe_col_len, ht, ht_min, ht_max, stream=None):
    """full_join (pointer API), kernels/join.hpp:314-355 (in place on flags)."""
    probe_full(probe_col, probe_val_out, probe_flags, probe_flags, ht, ht_min, ht_max, n=probe_col_len,
               stream=stream)


def aggregate_operation(a, flags, size, agg_res, stream=None):
    """aggregate_operation, kernels/aggregation.hpp:230-276 (accumulates)."""
    sum_flagged(a, flags, agg_res, n=size, stream=stream)


def group_by_aggregate(contents, agg_column, max, min, flags, col_len, col_num, results, agg_result,
                       result_flags, prod_ranges, agg_op="SUM", stream=None):
    """group_by_aggregate (pointer API), kernels/aggregation.hpp:346-409."""
    if len(contents) != col_num or len(results) != col_num:
        raise ValueError("col_num does not match the number of columns")
    groupby_dense(contents, list(min), list(max), agg_column, flags, results, agg_result, result_flags,
                  prod_ranges, n=col_len, stream=stream)


def count_true_flags(flags, length, stream=None):
    """count_true_flags, kernels/common.hpp:67-79 (synchronous)."""
    out = torch.zeros(1, dtype=torch.int64, device=flags.device)
    count_flags(flags, out, accumulate=False, n=length, stream=stream)
    return int(out.item())


def compress_sync(cols, flags, host_cols, n=None, stream=None):
    """Segment::compress_sync (models/models.hpp:990-1068): the rows whose
    device flag is set are copied from the device columns (torch int32) into
    the same rows of the host columns (numpy int32, C-contiguous, written in
    place); returns the number of rows synced."""
    n = _n(n, flags)
    k = len(cols)
    if len(host_cols) != k:
        raise ValueError("one host column per device column")
    dp = (ctypes.c_void_p * max(k, 1))(*[_dev(c, torch.int32, "column", n).value for c in cols])
    hp = (ctypes.c_void_p * max(k, 1))()
    for i, h in enumerate(host_cols):
        if h.dtype != "int32" or not h.flags["C_CONTIGUOUS"] or not h.flags["WRITEABLE"] or len(h) < n:
            raise ValueError("host column %d must be a writable contiguous int32 array of n rows" % i)
        hp[i] = h.ctypes.data
    cnt = ctypes.c_uint64()
    L.check(L.lib.sdb_compress_sync(dp, k, _dev(flags, torch.uint8, "flags", n), n, hp, ctypes.byref(cnt),
                                    _stream(stream)))
    return cnt.value


def bundle_exec(ops, n, stream=None):
    """KernelBundle::execute with fusion (models/execution.hpp:528-650):
    `ops` is a list of sycldb_amd._lib.Op (sdb_op) run in order over n rows,
    as one fused kernel when they allow it."""
    arr = (L.Op * max(len(ops), 1))(*ops)
    L.check(L.lib.sdb_bundle_exec(arr, len(ops), int(n), _stream(stream)))
