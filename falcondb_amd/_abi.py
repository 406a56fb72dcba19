"""ctypes declarations of include/falcon_gpu.h and the loader of the native
library.  The library is loaded from this package directory only; if it is
missing the import of any operator fails loudly (there is no CPU fallback)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
GPU_LIB_PATH = PKG / "libfalcon_gpu.so"

# status codes (FalconError variants in declaration order, then device errors)
OK = 0
ERROR_NAMES = {
    1: "ColumnNotFound", 2: "TypeMismatch", 3: "IndexOutOfBounds", 4: "ParseError",
    5: "Io", 6: "InvalidQuery", 7: "DivisionByZero", 8: "Overflow",
    9: "ConstraintViolation", 100: "Device", 101: "OutOfMemory", 102: "Unsupported",
}

E_INDEX_OUT_OF_BOUNDS, E_UNSUPPORTED = 3, 102

INT32, INT64, FLOAT64, UTF8 = 1, 2, 3, 4
EQ, NE, LT, LE, GT, GE = range(6)
P_CMP, P_AND, P_OR, P_IS_NULL, P_IS_NOT_NULL, P_IN, P_LIKE = range(7)
COUNT, SUM, AVG, MIN, MAX = range(5)
INNER, LEFT, SEMI, ANTI = range(4)
NO_ROW = 0xFFFFFFFF
MAX_GROUP_KEYS = 4  # FGPU_MAX_GROUP_KEYS
NULL_CODE = 0xFFFFFFFF


class ColumnDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", C.c_int32), ("values", C.c_void_p),
                ("validity", C.c_void_p), ("dict_len", C.c_int32),
                ("dict", C.POINTER(C.c_char_p))]


class Utf8Batches(C.Structure):
    _fields_ = [("n_batches", C.c_int32), ("batch_rows", C.POINTER(C.c_int64)),
                ("dict_len", C.POINTER(C.c_int32)), ("dicts", C.POINTER(C.POINTER(C.c_char_p)))]


class PredNode(C.Structure):
    _fields_ = [("kind", C.c_int32), ("col", C.c_int32), ("op", C.c_int32),
                ("lit_type", C.c_int32), ("lit_i64", C.c_int64), ("lit_f64", C.c_double),
                ("lit_str", C.c_char_p), ("n_list", C.c_int32), ("list", C.c_void_p)]


class Pred(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("nodes", C.POINTER(PredNode))]


class Agg(C.Structure):
    _fields_ = [("func", C.c_int32), ("col", C.c_int32), ("side", C.c_int32),
                ("alias", C.c_char_p)]


class GroupBySpec(C.Structure):
    _fields_ = [("key_col", C.c_int32), ("n_aggs", C.c_int32), ("aggs", C.POINTER(Agg))]


class JoinSpec(C.Structure):
    _fields_ = [("type", C.c_int32), ("probe_key", C.c_int32), ("build_key", C.c_int32),
                ("n_aggs", C.c_int32), ("aggs", C.POINTER(Agg))]


class SortKey(C.Structure):
    _fields_ = [("col", C.c_int32), ("desc", C.c_int32), ("nulls_last", C.c_int32)]


class ColumnView(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dtype", C.c_int32), ("values", C.c_void_p),
                ("valid", C.POINTER(C.c_uint8)), ("dict_len", C.c_int32),
                ("dict", C.POINTER(C.c_char_p))]


class BatchQuery(C.Structure):
    _fields_ = [("table", C.c_void_p), ("pred", C.c_void_p), ("spec", C.c_void_p)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("main_ms", C.c_float), ("build_ms", C.c_float),
                ("finalize_ms", C.c_float), ("bytes_main", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("table", C.c_void_p), ("pred", C.c_void_p),
                ("groupby", C.c_void_p), ("build", C.c_void_p), ("index", C.c_void_p),
                ("join", C.c_void_p), ("topk_col", C.c_int32), ("topk_desc", C.c_int32),
                ("topk_k", C.c_int32), ("alias", C.c_char_p)]


class RequestStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("rows_in", C.c_int64), ("reserved_bytes", C.c_int64), ("slot", C.c_int32),
                ("shared", C.c_int32), ("retries", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("shared_launches", C.c_int64), ("shared_requests", C.c_int64),
                ("retries", C.c_int64), ("trims", C.c_int64), ("peak_held_bytes", C.c_int64),
                ("peak_running", C.c_int32), ("n_slots", C.c_int32), ("budget_bytes", C.c_int64)]


REQ_COUNT, REQ_GROUPBY, REQ_JOIN, REQ_TOPK = range(4)


# Every symbol include/falcon_gpu.h declares (the CPU test checks the library
# exports each of them).
DECLARED = {
    "fgpu_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "fgpu_close": (None, [C.c_void_p]),
    "fgpu_last_error": (C.c_char_p, []),
    "fgpu_abi_version": (C.c_int, []),
    "fgpu_stream": (C.c_void_p, [C.c_void_p]),
    "fgpu_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_last_timing": (C.c_int, [C.c_void_p, C.POINTER(Timing)]),
    "fgpu_scratch_bytes": (C.c_int64, [C.c_void_p]),
    "fgpu_table_upload": (C.c_int, [C.c_void_p, C.POINTER(ColumnDesc), C.c_int, C.c_int64,
                                    C.c_int32, C.POINTER(C.c_void_p)]),
    "fgpu_table_upload_batches": (C.c_int, [C.c_void_p, C.POINTER(ColumnDesc), C.POINTER(Utf8Batches), C.c_int,
                                            C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "fgpu_table_free": (None, [C.c_void_p]),
    "fgpu_table_rows": (C.c_int64, [C.c_void_p]),
    "fgpu_table_device_bytes": (C.c_int64, [C.c_void_p]),
    "fgpu_table_column_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]),
    "fgpu_filter_count": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.POINTER(C.c_int64)]),
    "fgpu_filter_mask": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.c_void_p]),
    "fgpu_filter_select": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred),
                                     C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_void_p)]),
    "fgpu_groupby": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.POINTER(GroupBySpec),
                               C.POINTER(C.c_void_p)]),
    "fgpu_groupby_keys": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.POINTER(C.c_int32), C.c_int32,
                                    C.POINTER(GroupBySpec), C.POINTER(C.c_void_p)]),
    "fgpu_groupby_static": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.POINTER(GroupBySpec),
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.c_int64, C.POINTER(C.c_int64)]),
    "fgpu_join_materialize": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.c_void_p, C.c_void_p,
                                        C.POINTER(JoinSpec), C.POINTER(C.c_int32), C.c_int,
                                        C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_void_p)]),
    "fgpu_sort_limit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.POINTER(SortKey), C.c_int,
                                  C.c_int64, C.POINTER(C.c_int32), C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "fgpu_table_cols": (C.c_int, [C.c_void_p]),
    "fgpu_table_column_info": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ColumnView)]),
    "fgpu_join_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "fgpu_join_index_free": (None, [C.c_void_p]),
    "fgpu_join_index_device_bytes": (C.c_int64, [C.c_void_p]),
    "fgpu_join_agg": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.c_void_p, C.c_void_p,
                                C.POINTER(JoinSpec), C.POINTER(C.c_void_p)]),
    "fgpu_join_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.c_void_p, C.c_void_p,
                                  C.POINTER(JoinSpec), C.c_void_p, C.c_void_p, C.c_int64,
                                  C.POINTER(C.c_int64)]),
    "fgpu_topk": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Pred), C.c_int32, C.c_int, C.c_int32,
                            C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "fgpu_batch_prepare": (C.c_int, [C.c_void_p, C.POINTER(BatchQuery), C.c_int, C.POINTER(C.c_void_p)]),
    "fgpu_batch_tune": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "fgpu_batch_info": (C.c_char_p, [C.c_void_p]),
    "fgpu_batch_fallbacks": (C.c_int, [C.c_void_p]),
    "fgpu_batch_free": (None, [C.c_void_p]),
    "fgpu_server_open": (C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_void_p)]),
    "fgpu_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "fgpu_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p),
                                   C.POINTER(RequestStats)]),
    "fgpu_server_pause": (C.c_int, [C.c_void_p, C.c_int]),
    "fgpu_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "fgpu_server_close": (None, [C.c_void_p]),
    "fgpu_result_rows": (C.c_int64, [C.c_void_p]),
    "fgpu_result_cols": (C.c_int, [C.c_void_p]),
    "fgpu_result_column": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ColumnView)]),
    "fgpu_result_free": (None, [C.c_void_p]),
}

# Exported helpers beyond the header's operator set.
EXTRA = {
    "fgpu_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "fgpu_table_float_range": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "fgpu_host_exact_sum": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "fgpu_last_plan": (C.c_char_p, [C.c_void_p]),
}

_lib = None


def load(path: Path | None = None):
    """Loads libfalcon_gpu.so (built by `python -m falcondb_amd.build`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else GPU_LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is missing: build it with `python -m falcondb_amd.build` "
            "(the operators have no CPU fallback)")
    lib = C.CDLL(str(p))
    for table in (DECLARED, EXTRA):
        for name, (res, args) in table.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
