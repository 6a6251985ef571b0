"""ctypes/numpy front of the CPU oracle (oracle/sdb_oracle.c).

TEST INFRASTRUCTURE ONLY: this module is the checker that the HIP product is
compared against.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product
(``sycldb_amd``) never does.

Each wrapper takes numpy arrays (int32 columns, uint8 flags, uint64
accumulators), checks shapes/dtypes, and calls the C restatement.  The C file
cites, per function, the reference functor it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

EQ, NE, LT, LE, GT, GE = range(6)          # kernels/selection.hpp:7-15
NONE, AND, OR = range(3)                   # kernels/selection.hpp:16-21
MUL, DIV, ADD, SUB = range(4)              # kernels/aggregation.hpp:11-17


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return ctypes.CDLL(_LIB_PATH)


_lib = _load()

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_U64 = ctypes.c_uint64
_INT = ctypes.c_int

_sigs = {
    "orc_select_lit": [_P, _P, _P, _I64, _INT, _INT, _I32],
    "orc_select_col": [_P, _P, _P, _P, _I64, _INT, _INT],
    "orc_logical": [_P, _P, _P, _I64, _INT],
    "orc_fill_i32": [_P, _I32, _I64],
    "orc_binop_cc": [_P, _P, _P, _P, _I64, _INT],
    "orc_binop_lc": [_P, _I32, _P, _P, _I64, _INT],
    "orc_binop_cl": [_P, _P, _I32, _P, _I64, _INT],
    "orc_build_keys": [_P, _P, _I64, _P, _I32, _I32],
    "orc_build_keyvals": [_P, _P, _P, _I64, _P, _I32, _I32],
    "orc_probe_semi": [_P, _P, _P, _I64, _P, _I32, _I32],
    "orc_probe_full": [_P, _P, _P, _P, _I64, _P, _I32, _I32],
    "orc_sum_flagged": [_P, _P, _I64, _P],
    "orc_groupby_dense": [_P, _INT, _P, _P, _P, _P, _I64, _P, _P, _P, _U64],
    "orc_seen_to_flags": [_P, _P, _U64],
    "orc_count_flags": [_P, _I64],
    "orc_compact": [_P, _I64, _P],
    "orc_minmax": [_P, _I64, _P, _P],
    "orc_gen_uniform": [_P, _I64, _U64, _U64, _I64, _I32, _I32],
    "orc_gen_child": [_P, _P, _I64, _U64, _U64, _I64, _I32, _I32, _I32],
    "orc_gen_lookup": [_P, _I64, _U64, _U64, _I64, _P, _I32],
    "orc_gen_dates": [_P, _P, _P, _I32],
    "orc_num_threads": [],
    "orc_host_copy_gbps": [_I64],
}
for _name, _args in _sigs.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = None
_lib.orc_count_flags.restype = _U64
_lib.orc_compact.restype = _I64
_lib.orc_num_threads.restype = _INT
_lib.orc_host_copy_gbps.restype = ctypes.c_double


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    assert isinstance(a, np.ndarray) and a.dtype == np.int32 and a.flags.c_contiguous, \
        "expected contiguous int32 array"
    return a


def _u8(a):
    assert isinstance(a, np.ndarray) and a.dtype == np.uint8 and a.flags.c_contiguous, \
        "expected contiguous uint8 flags"
    return a


def num_threads():
    return _lib.orc_num_threads()


def host_copy_gbps(nbytes=1 << 30):
    """Host memory bandwidth (parallel copy, read + write), GB/s."""
    return _lib.orc_host_copy_gbps(int(nbytes))


def select_lit(col, in_flags, out_flags, cmp, logic, value):
    n = len(_i32(col))
    assert len(_u8(in_flags)) >= n and len(_u8(out_flags)) >= n
    _lib.orc_select_lit(_ptr(col), _ptr(in_flags), _ptr(out_flags), n, cmp, logic, int(value))


def select_col(a, b, in_flags, out_flags, cmp, logic):
    n = len(_i32(a))
    assert len(_i32(b)) >= n and len(_u8(in_flags)) >= n and len(_u8(out_flags)) >= n
    _lib.orc_select_col(_ptr(a), _ptr(b), _ptr(in_flags), _ptr(out_flags), n, cmp, logic)


def logical(in_flags, local_flags, out_flags, logic):
    n = len(_u8(in_flags))
    assert len(_u8(local_flags)) >= n and len(_u8(out_flags)) >= n
    _lib.orc_logical(_ptr(in_flags), _ptr(local_flags), _ptr(out_flags), n, logic)


def fill_i32(dst, value):
    _lib.orc_fill_i32(_ptr(_i32(dst)), int(value), len(dst))


def binop_cc(result, a, b, flags, op):
    n = len(_i32(a))
    _i32(result), _i32(b), _u8(flags)
    _lib.orc_binop_cc(_ptr(result), _ptr(a), _ptr(b), _ptr(flags), n, op)


def binop_lc(result, a, b, flags, op):
    n = len(_i32(b))
    _i32(result), _u8(flags)
    _lib.orc_binop_lc(_ptr(result), int(a), _ptr(b), _ptr(flags), n, op)


def binop_cl(result, a, b, flags, op):
    n = len(_i32(a))
    _i32(result), _u8(flags)
    _lib.orc_binop_cl(_ptr(result), _ptr(a), int(b), _ptr(flags), n, op)


def build_keys(col, flags, ht, ht_len, ht_min):
    _i32(col), _u8(flags), _u8(ht)
    assert len(ht) >= ht_len
    _lib.orc_build_keys(_ptr(col), _ptr(flags), len(col), _ptr(ht), ht_len, ht_min)


def build_keyvals(col, vals, flags, ht, ht_len, ht_min):
    _i32(col), _i32(vals), _u8(flags), _i32(ht)
    assert len(ht) >= 2 * ht_len
    _lib.orc_build_keyvals(_ptr(col), _ptr(vals), _ptr(flags), len(col), _ptr(ht), ht_len, ht_min)


def probe_semi(probe, in_flags, out_flags, ht, ht_min, ht_max):
    _i32(probe), _u8(in_flags), _u8(out_flags), _u8(ht)
    _lib.orc_probe_semi(_ptr(probe), _ptr(in_flags), _ptr(out_flags), len(probe), _ptr(ht),
                        ht_min, ht_max)


def probe_full(probe, val_out, in_flags, out_flags, ht, ht_min, ht_max):
    _i32(probe), _i32(val_out), _u8(in_flags), _u8(out_flags), _i32(ht)
    _lib.orc_probe_full(_ptr(probe), _ptr(val_out), _ptr(in_flags), _ptr(out_flags), len(probe),
                        _ptr(ht), ht_min, ht_max)


def sum_flagged(col, flags, acc):
    """acc: numpy uint64 array of length 1, accumulated into."""
    _i32(col), _u8(flags)
    assert acc.dtype == np.uint64
    _lib.orc_sum_flagged(_ptr(col), _ptr(flags), len(col), _ptr(acc))


def groupby_dense(keys, mins, maxs, agg, flags, out_keys, sums, seen, prod):
    k = len(keys)
    keys = [_i32(c) for c in keys]
    key_ptrs = (ctypes.c_void_p * k)(*[c.ctypes.data for c in keys])
    out_ptrs = (ctypes.c_void_p * k)(*[_i32(c).ctypes.data for c in out_keys])
    mins_a = np.ascontiguousarray(mins, dtype=np.int32)
    maxs_a = np.ascontiguousarray(maxs, dtype=np.int32)
    assert sums.dtype == np.uint64 and seen.dtype == np.uint32
    assert len(sums) >= prod and len(seen) >= prod
    _lib.orc_groupby_dense(key_ptrs, k, _ptr(mins_a), _ptr(maxs_a), _ptr(_i32(agg)), _ptr(_u8(flags)),
                           len(agg), out_ptrs, _ptr(sums), _ptr(seen), int(prod))


def seen_to_flags(seen, flags):
    _lib.orc_seen_to_flags(_ptr(seen), _ptr(_u8(flags)), len(seen))


def count_flags(flags):
    return int(_lib.orc_count_flags(_ptr(_u8(flags)), len(flags)))


def compact(flags):
    out = np.empty(len(flags), dtype=np.int32)
    c = _lib.orc_compact(_ptr(_u8(flags)), len(flags), _ptr(out))
    return out[:c].copy()


def minmax(col):
    mn = np.zeros(1, np.int32)
    mx = np.zeros(1, np.int32)
    _lib.orc_minmax(_ptr(_i32(col)), len(col), _ptr(mn), _ptr(mx))
    return int(mn[0]), int(mx[0])


# ---- synthetic data generator (same data spec as the product's device one) ----

def gen_uniform(n, seed, stream, lo, hi, first_row=0):
    out = np.empty(n, dtype=np.int32)
    _lib.orc_gen_uniform(_ptr(out), n, seed, stream, first_row, lo, hi)
    return out


def gen_child(base, seed, stream, mult, lo, hi, first_row=0):
    base = _i32(np.ascontiguousarray(base, dtype=np.int32))
    out = np.empty(len(base), dtype=np.int32)
    _lib.orc_gen_child(_ptr(out), _ptr(base), len(base), seed, stream, first_row, mult, lo, hi)
    return out


def gen_lookup(n, seed, stream, table, first_row=0):
    table = _i32(np.ascontiguousarray(table, dtype=np.int32))
    out = np.empty(n, dtype=np.int32)
    _lib.orc_gen_lookup(_ptr(out), n, seed, stream, first_row, _ptr(table), len(table))
    return out


def gen_dates(ndays):
    dk = np.empty(ndays, np.int32)
    yr = np.empty(ndays, np.int32)
    ym = np.empty(ndays, np.int32)
    _lib.orc_gen_dates(_ptr(dk), _ptr(yr), _ptr(ym), ndays)
    return dk, yr, ym
