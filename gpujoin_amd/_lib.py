"""ctypes binding of libgj.so (include/gj.h and include/gj_runtime.h).

The shared library is built in-tree by ``make -C gpujoin_amd``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgj.so")

NOT_FOUND = 0xFFFFFFFF

# gj_status
GJ_OK = 0
GJ_ERR_INVALID = -1
GJ_ERR_UNSUPPORTED = -2
GJ_ERR_HIP = -3
GJ_ERR_NOMEM = -4
GJ_ERR_OVERFLOW = -5
GJ_ERR_INTERNAL = -6

# gj_index_kind
BINARY_SEARCH, BTREE, HARMONIA, RADIX_SPLINE, LEARNED_PLR, LEARNED_MLP, NO_OP, HASH = range(8)
INDEX_NAMES = {
    "binary_search": BINARY_SEARCH, "btree": BTREE, "harmonia": HARMONIA, "radix_spline": RADIX_SPLINE,
    "learned_plr": LEARNED_PLR, "learned_mlp": LEARNED_MLP, "no_op": NO_OP, "hash": HASH,
}
# gj_approach
PLAIN, BWS, PARTITIONING, HJ = range(4)
APPROACH_NAMES = {"plain": PLAIN, "bws": BWS, "partitioning": PARTITIONING, "hj": HJ}
# gj_batch_mode
BATCH_LAUNCH, BATCH_GRAPH, BATCH_PERSISTENT = range(3)
# generators
KEYS_DENSE, KEYS_UNIFORM, KEYS_LOGNORMAL, KEYS_CLUSTERED = range(4)
PROBES_UNIFORM, PROBES_ZIPF = range(2)
# gj_index_params.learned_flags / learned_layout
LEARNED_UNTRAINED = 0x1
LAYOUT_AUTO, LAYOUT_COLUMN, LAYOUT_LINES = range(3)
# gj_index_params.placement
PLACE_DEVICE, PLACE_HOST = range(2)
# gj_agg_strategy
AGG_LANE_REFILL, AGG_PLAIN, AGG_PARTITIONED, AGG_SORTED, AGG_LOOKUP_FIRST = range(5)
AGG_STRATEGIES = {"lane_refill": AGG_LANE_REFILL, "plain": AGG_PLAIN, "partitioned": AGG_PARTITIONED,
                  "sorted": AGG_SORTED, "lookup_first": AGG_LOOKUP_FIRST}


class IndexParams(C.Structure):
    _fields_ = [
        ("rs_radix_bits", C.c_uint32), ("rs_max_error", C.c_uint32), ("harmonia_group_lanes", C.c_uint32),
        ("btree_lds_bytes", C.c_uint32), ("learned_leaf_bits", C.c_uint32), ("learned_max_error", C.c_uint32),
        ("mlp_hidden", C.c_uint32), ("mlp_steps", C.c_uint32), ("mlp_batch", C.c_uint32), ("mlp_lr", C.c_float),
        ("seed", C.c_uint64), ("learned_root_segments", C.c_uint32), ("learned_flags", C.c_uint32),
        ("learned_layout", C.c_uint32), ("learned_line_keys", C.c_uint32), ("rs_lds_bits", C.c_uint32),
        ("placement", C.c_uint32), ("btree_node_bytes", C.c_uint32),
    ]


class LookupParams(C.Structure):
    _fields_ = [
        ("approach", C.c_int), ("block_size", C.c_uint32), ("grid_blocks", C.c_uint32), ("radix_bits", C.c_uint32),
        ("passes", C.c_uint32), ("ignore_bits", C.c_uint32), ("window", C.c_uint64),
    ]


class IndexStats(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("key_bytes", C.c_int), ("n", C.c_uint64), ("memory_bytes", C.c_uint64),
        ("depth", C.c_uint32), ("lds_levels", C.c_uint32), ("spline_points", C.c_uint64), ("leaves", C.c_uint64),
        ("fallback_leaves", C.c_uint64), ("fallback_keys", C.c_uint64), ("max_window", C.c_uint32),
        ("avg_window", C.c_double), ("train_loss", C.c_float), ("root_segments", C.c_uint32),
        ("split_leaves", C.c_uint32), ("dropped_tuples", C.c_uint64), ("layout", C.c_uint32),
        ("max_lines", C.c_uint32), ("lines", C.c_uint64), ("avg_lines", C.c_double), ("root_trained", C.c_uint32),
        ("placement", C.c_uint32),
    ]


class PartitionParams(C.Structure):
    _fields_ = [("radix_bits", C.c_uint32), ("ignore_bits", C.c_uint32), ("passes", C.c_uint32)]


class Partitioned(C.Structure):
    _fields_ = [
        ("d_keys", C.c_void_p), ("d_payloads", C.c_void_p), ("d_offsets", C.c_void_p), ("d_counts", C.c_void_p),
        ("capacity", C.c_uint64), ("d_dropped", C.c_void_p),
    ]


class FactBatch(C.Structure):
    _fields_ = [
        ("d_fkey", C.c_void_p), ("d_date", C.c_void_p), ("d_price", C.c_void_p), ("d_discount", C.c_void_p),
        ("rows", C.c_uint64),
    ]


class AggParams(C.Structure):
    _fields_ = [("date_lo", C.c_int32), ("date_hi", C.c_int32), ("d_flag", C.c_void_p), ("strategy", C.c_uint32),
                ("radix_bits", C.c_uint32)]


class AggResult(C.Structure):
    _fields_ = [("numerator", C.c_int64), ("denominator", C.c_int64), ("matches", C.c_uint64), ("pad", C.c_uint64)]


class ProbeBatch(C.Structure):
    _fields_ = [("d_probe_keys", C.c_void_p), ("d_rowids", C.c_void_p), ("m", C.c_uint64)]


_P = C.c_void_p
_U64 = C.c_uint64
_U32 = C.c_uint32
_I = C.c_int

# name -> (restype, argtypes); every symbol declared in include/gj.h and include/gj_runtime.h
SIGNATURES = {
    # gj.h
    "gj_index_params_default": (None, [C.POINTER(IndexParams)]),
    "gj_lookup_params_default": (None, [C.POINTER(LookupParams)]),
    "gj_index_build": (_I, [_I, _P, _P, _U64, _I, C.POINTER(IndexParams), _P, C.POINTER(_P)]),
    "gj_index_lookup": (_I, [_P, _P, _U64, _P, C.POINTER(LookupParams), _P]),
    "gj_index_memory": (_U64, [_P]),
    "gj_index_get_stats": (_I, [_P, C.POINTER(IndexStats)]),
    "gj_index_destroy": (None, [_P]),
    "gj_parse_index_type": (_I, [C.c_char_p, C.POINTER(_I)]),
    "gj_index_type_name": (C.c_char_p, [_I]),
    "gj_partition_capacity": (_U64, [_U64, _I, _U32]),
    "gj_partition_scratch_bytes": (_U64, [_U64, _I, C.POINTER(PartitionParams)]),
    "gj_partition": (_I, [_P, _U64, _I, C.POINTER(PartitionParams), C.POINTER(Partitioned), _P, _U64, _P]),
    "gj_join_aggregate": (_I, [_P, C.POINTER(FactBatch), C.POINTER(AggParams), _P, _P]),
    "gj_join_aggregate_noalloc": (_I, [_P, C.POINTER(FactBatch), C.POINTER(AggParams), _P, _P]),
    "gj_promo_revenue": (C.c_int64, [C.c_int64, C.c_int64]),
    "gj_lookup_plan_create": (_I, [_P, C.POINTER(ProbeBatch), _U32, _I, _P, C.POINTER(_P)]),
    "gj_aggregate_plan_create": (_I, [_P, C.POINTER(FactBatch), _U32, C.POINTER(AggParams), _P, _I, _P,
                                      C.POINTER(_P)]),
    "gj_batch_plan_run": (_I, [_P, _P]),
    "gj_batch_plan_destroy": (None, [_P]),
    "gj_autotune_lookup": (_I, [_P, _P, _U64, _P, C.POINTER(LookupParams), C.POINTER(C.c_double), _P]),
    "gj_autotune_index": (_I, [_I, _P, _P, _U64, _I, C.POINTER(IndexParams), _P, _U64, _P, C.POINTER(IndexParams),
                               C.POINTER(C.c_double), _P]),
    "gj_mlp_loss_grad": (_I, [_P, _U32, _P, _P, _U64, _P, _P, _P]),
    "gj_mlp_forward": (_I, [_P, _U32, _P, _U64, _P, _P]),
    "gj_gather": (_I, [_P, _U64, _P, _U64, _P, _U32, _U32, _P]),
    "gj_pointer_chase": (_I, [_P, _U64, _U32, _U32, _P, _P, _P]),
    "gj_stream_copy": (_I, [_P, _P, _U64, _P]),
    "gj_stream_read": (_I, [_P, _U64, _P, _P]),
    "gj_last_error": (C.c_char_p, []),
    "gj_version": (C.c_char_p, []),
    # gj_runtime.h
    "gj_device_count": (_I, [C.POINTER(_I)]),
    "gj_device_info": (_I, [C.c_char_p, C.c_size_t, C.POINTER(_I), C.POINTER(_I)]),
    "gj_malloc": (_I, [C.POINTER(_P), _U64]),
    "gj_free": (_I, [_P]),
    "gj_host_malloc": (_I, [C.POINTER(_P), _U64]),
    "gj_host_free": (_I, [_P]),
    "gj_memcpy_htod": (_I, [_P, _P, _U64, _P]),
    "gj_memcpy_dtoh": (_I, [_P, _P, _U64, _P]),
    "gj_memcpy_dtod": (_I, [_P, _P, _U64, _P]),
    "gj_memset": (_I, [_P, _I, _U64, _P]),
    "gj_stream_create": (_I, [C.POINTER(_P)]),
    "gj_stream_destroy": (_I, [_P]),
    "gj_stream_synchronize": (_I, [_P]),
    "gj_device_synchronize": (_I, []),
    "gj_event_create": (_I, [C.POINTER(_P)]),
    "gj_event_destroy": (_I, [_P]),
    "gj_event_record": (_I, [_P, _P]),
    "gj_event_elapsed_ms": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "gj_gen_build_keys": (_I, [_P, _U64, _I, _I, _U32, _U64, _I]),
    "gj_gen_probes": (_I, [_P, _U64, _I, _P, _U64, _I, C.c_double, C.c_double, _U32, _U64, _P, _I]),
    "gj_gen_dimension": (_I, [_P, _P, _U64, _U64, _I]),
    "gj_gen_fact": (_I, [_P, _P, _P, _P, _U64, _U64, C.c_int32, C.c_int64, C.c_int64, C.c_int32, _U64, _I]),
    "gj_sort_dimension": (_I, [_P, _P, _U64, _P, _P]),
    "gj_sort_keys": (_I, [_P, _U64, _I, _I]),
    "gj_rng_u64": (_U64, [_U64, _U64]),
    "gj_zipf_trial": (_U64, [_U64, C.c_double, C.c_double]),
}

_lib = None


class GjError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gj status {status}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Loads libgj.so (once).  Raises when the library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C gpujoin_amd` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by exported_symbols(); calls fail loudly below
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> dict:
    lib = load()
    return {name: hasattr(lib, name) for name in SIGNATURES}


def check(status: int) -> None:
    if status != GJ_OK:
        raise GjError(status, load().gj_last_error().decode(errors="replace"))


def call(name: str, *args):
    fn = getattr(load(), name, None)
    if fn is None:
        raise GjError(GJ_ERR_UNSUPPORTED, f"{name} is not exported by {LIB_PATH}")
    return fn(*args)
