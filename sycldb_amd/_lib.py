"""ctypes binding of libsycldb_hip.so (the C ABI declared in include/sycldb.h).

The library is the product: there is no fallback.  Importing this module
raises ImportError when the built library is missing.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsycldb_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libsycldb_hip.so is not built (expected at %s); run __graft_entry__.build() "
        "or `python -m sycldb_amd.build`" % LIB_PATH)



def _share_hip_runtime():
    """Keeps one HIP runtime per process.

    PyTorch for ROCm ships its own libamdhip64 (SONAME libamdhip64.so.7) and
    HSA runtime, which its libraries request as "libamdhip64.so".  If this
    library were loaded first it would bring in the system ROCm runtime under
    the .so.7 name, and a later `import torch` would then load a second HIP and
    HSA runtime whose device initialisation fails.  Loading torch's runtime
    first, globally, makes this library bind to it (same SONAME) and torch
    reuse it (same file), whichever of the two is imported first.  Without
    torch installed the system runtime is used.
    """
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return None
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


_hip_runtime = _share_hip_runtime()
lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
U64 = ctypes.c_uint64
INT = ctypes.c_int
CSTR = ctypes.c_char_p
DBL = ctypes.c_double

SDB_OK, SDB_ERR_INVALID_ARG, SDB_ERR_HIP, SDB_ERR_OOM, SDB_ERR_UNSUPPORTED, SDB_ERR_RUNTIME = range(6)
EQ, NE, LT, LE, GT, GE = range(6)
NONE, AND, OR = range(3)
MUL, DIV, ADD, SUB = range(4)
CODEGEN_AUTO, CODEGEN_INTERPRET, CODEGEN_COMPILE = range(3)


class EngineOptions(ctypes.Structure):
    _fields_ = [("device_budget_bytes", ctypes.c_uint64),
                ("segment_size", ctypes.c_uint64),
                ("disable_fusion", ctypes.c_int),
                ("device", ctypes.c_int)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int),
                ("memory_budget_bytes", ctypes.c_uint64),
                ("serial", ctypes.c_int),
                ("single_worker", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("latency_ms", ctypes.c_double),
                ("queue_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double),
                ("batch_size", ctypes.c_int)]


# sdb_op_kind (KernelType order)
(OP_EMPTY, OP_LOGICAL, OP_SELECT_COL, OP_SELECT_LIT, OP_FILL, OP_BINOP_CC, OP_BINOP_LC, OP_BINOP_CL,
 OP_BUILD_KEYS, OP_PROBE_SEMI, OP_BUILD_KEYVALS, OP_PROBE_FULL, OP_SUM, OP_GROUPBY) = range(14)
OP_MAX_KEYS = 8
BUNDLE_MAX_OPS = 64


class Op(ctypes.Structure):
    """sdb_op: one operator of a bundle (sdb_bundle_exec)."""
    _fields_ = [("kind", ctypes.c_int32), ("cmp", ctypes.c_int32), ("logic", ctypes.c_int32),
                ("binop", ctypes.c_int32), ("ht_min", ctypes.c_int32), ("ht_max", ctypes.c_int32),
                ("ht_len", ctypes.c_int32), ("value", ctypes.c_int32),
                ("col", P), ("col2", P), ("res", P), ("ht", P),
                ("in_flags", P), ("out_flags", P), ("local_flags", P), ("acc", P),
                ("nkeys", ctypes.c_int32), ("mins", ctypes.c_int32 * OP_MAX_KEYS),
                ("maxs", ctypes.c_int32 * OP_MAX_KEYS), ("keys", P * OP_MAX_KEYS), ("out_keys", P * OP_MAX_KEYS),
                ("sums", P), ("seen", P), ("prod", ctypes.c_uint64)]


# name: (restype, argtypes)
SIGNATURES = {
    "sdb_last_error": (CSTR, []),
    "sdb_version": (CSTR, []),
    "sdb_comp_op_from_string": (INT, [CSTR]),
    "sdb_logical_op_from_string": (INT, [CSTR]),
    "sdb_binop_from_string": (INT, [CSTR]),
    "sdb_select_lit": (INT, [P, P, P, I64, INT, INT, I32, P]),
    "sdb_select_col": (INT, [P, P, P, P, I64, INT, INT, P]),
    "sdb_logical": (INT, [P, P, P, I64, INT, P]),
    "sdb_fill_i32": (INT, [P, I32, I64, P]),
    "sdb_binop_cc": (INT, [P, P, P, P, I64, INT, P]),
    "sdb_binop_lc": (INT, [P, I32, P, P, I64, INT, P]),
    "sdb_binop_cl": (INT, [P, P, I32, P, I64, INT, P]),
    "sdb_build_keys": (INT, [P, P, I64, P, I32, I32, P]),
    "sdb_build_keyvals": (INT, [P, P, P, I64, P, I32, I32, P]),
    "sdb_probe_semi": (INT, [P, P, P, I64, P, I32, I32, P]),
    "sdb_probe_full": (INT, [P, P, P, P, I64, P, I32, I32, P]),
    "sdb_sum_flagged": (INT, [P, P, I64, P, P]),
    "sdb_groupby_dense": (INT, [P, INT, P, P, P, P, I64, P, P, P, U64, P]),
    "sdb_seen_to_flags": (INT, [P, P, U64, P]),
    "sdb_count_flags": (INT, [P, I64, P, INT, P]),
    "sdb_compact": (INT, [P, I64, P, P, P]),
    "sdb_compact_noalloc": (INT, [P, I64, P, P, P]),
    "sdb_bundle_exec": (INT, [ctypes.POINTER(Op), INT, I64, P]),
    "sdb_gather_i32": (INT, [P, P, I64, P, P]),
    "sdb_compress_sync": (INT, [ctypes.POINTER(P), INT, P, I64, ctypes.POINTER(P), ctypes.POINTER(U64), P]),
    "sdb_minmax_i32": (INT, [P, I64, P, P, P]),
    "sdb_gen_uniform": (INT, [P, I64, U64, U64, I64, I32, I32, P]),
    "sdb_gen_child": (INT, [P, P, I64, U64, U64, I64, I32, I32, I32, P]),
    "sdb_gen_lookup": (INT, [P, I64, U64, U64, I64, P, I32, P]),
    "sdb_gen_rowid": (INT, [P, I64, I64, P]),
    "sdb_engine_create": (INT, [ctypes.POINTER(EngineOptions), ctypes.POINTER(P)]),
    "sdb_engine_destroy": (None, [P]),
    "sdb_engine_stream": (P, [P]),
    "sdb_engine_table_create": (INT, [P, CSTR, U64]),
    "sdb_engine_column_alloc": (INT, [P, CSTR, INT, ctypes.POINTER(P)]),
    "sdb_engine_column_upload": (INT, [P, CSTR, INT, P]),
    "sdb_engine_column_download": (INT, [P, CSTR, INT, P]),
    "sdb_engine_table_load_dir": (INT, [P, CSTR, CSTR]),
    "sdb_engine_table_finalize": (INT, [P, CSTR]),
    "sdb_engine_execute": (INT, [P, CSTR, ctypes.POINTER(P)]),
    "sdb_engine_last_timing": (INT, [P, ctypes.POINTER(DBL), ctypes.POINTER(DBL)]),
    "sdb_engine_reset": (INT, [P]),
    "sdb_engine_last_launches": (INT, [P, ctypes.POINTER(I64)]),
    "sdb_device_alloc": (INT, [ctypes.POINTER(P), U64]),
    "sdb_device_free": (INT, [P]),
    "sdb_copy_to_device": (INT, [P, P, U64, P]),
    "sdb_copy_to_host": (INT, [P, P, U64, P]),
    "sdb_stream_synchronize": (INT, [P]),
    "sdb_engine_execute_batch": (INT, [P, ctypes.POINTER(ctypes.c_char_p), INT, ctypes.POINTER(P),
                                       ctypes.POINTER(INT)]),
    "sdb_server_create": (INT, [ctypes.POINTER(P), INT, ctypes.POINTER(ServerOptions), ctypes.POINTER(P)]),
    "sdb_server_submit": (INT, [P, INT, CSTR, ctypes.POINTER(I64)]),
    "sdb_server_wait": (INT, [P, I64, ctypes.POINTER(P), ctypes.POINTER(RequestInfo)]),
    "sdb_server_stats": (INT, [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "sdb_server_destroy": (None, [P]),
    "sdb_engine_set_fusion": (INT, [P, INT]),
    "sdb_engine_set_codegen": (INT, [P, INT, INT]),
    "sdb_codegen_available": (INT, [CSTR, INT]),
    "sdb_codegen_kernels_compiled": (INT, []),
    "sdb_codegen_tuning_report": (INT, [ctypes.c_char_p, INT]),
    "sdb_codegen_check": (INT, [CSTR, ctypes.POINTER(INT)]),
    "sdb_codegen_wait": (INT, [ctypes.c_double]),
    "sdb_codegen_pending": (INT, []),
    "sdb_codegen_shutdown": (INT, []),
    "sdb_set_batch_graphs": (INT, [INT]),
    "sdb_batch_graph_stats": (INT, [ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "sdb_engine_set_profiling": (INT, [P, INT]),
    "sdb_engine_set_profile_table": (INT, [P, CSTR]),
    "sdb_engine_profile_reset": (INT, [P]),
    "sdb_engine_profile_count": (INT, [P]),
    "sdb_engine_profile_get": (INT, [P, INT, CSTR, INT, ctypes.POINTER(I64), ctypes.POINTER(I64),
                                      ctypes.POINTER(DBL)]),
    "sdb_result_nrows": (U64, [P]),
    "sdb_result_ncols": (INT, [P]),
    "sdb_result_col_is_u64": (INT, [P, INT]),
    "sdb_result_col_copy": (INT, [P, INT, P]),
    "sdb_result_flags_copy": (INT, [P, P]),
    "sdb_result_selected": (U64, [P]),
    "sdb_result_selected_rows": (INT, [P, P]),
    "sdb_result_col_copy_selected": (INT, [P, INT, P]),
    "sdb_result_write": (INT, [P, CSTR]),
    "sdb_result_destroy": (None, [P]),
    "sdb_plan_dag_order": (INT, [CSTR, P, INT, ctypes.POINTER(INT)]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

# The background kernel compiler stops while the interpreter exits, before
# the process tears down the compiler libraries a compilation in progress
# still uses (a no-op when nothing was compiled in the background).
atexit.register(lib.sdb_codegen_shutdown)


class SdbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sycldb error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    """Raise the Python counterpart of the C++ exception the reference would
    throw: ValueError for invalid_argument, MemoryError for bad_alloc."""
    if rc == SDB_OK:
        return
    msg = (lib.sdb_last_error() or b"").decode(errors="replace")
    if rc == SDB_ERR_INVALID_ARG:
        raise ValueError(msg)
    if rc == SDB_ERR_OOM:
        raise MemoryError(msg)
    raise SdbError(rc, msg)
