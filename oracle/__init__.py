"""ctypes binding of the oracle (oracle/oracle.c -> oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference algorithms that
the GPU path is checked against.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")
REF_ZIPF_PATH = os.path.join(_HERE, "_ref", "libref_zipf.so")
NOT_FOUND = 0xFFFFFFFF

_lib = None

_P = C.c_void_p
_U64 = C.c_uint64
_U32 = C.c_uint32
_I = C.c_int
_D = C.c_double

_SIGS = {
    "orc_lower_bound": (_U64, [_P, _U64, _I, _U64]),
    "orc_branchy_binary_search": (_U64, [_P, _U64, _I, _U64]),
    "orc_lookup_lower_bound": (None, [_P, _U64, _I, _P, _U64, _P, _I]),
    "orc_checksum": (_U64, [_P, _U64]),
    "orc_btree_capacities": (None, [_I, _I, C.POINTER(_U32), C.POINTER(_U32)]),
    "orc_btree_build": (_P, [_P, _U64, _I, _I]),
    "orc_btree_info": (None, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U64)]),
    "orc_btree_lookup": (None, [_P, _P, _U64, _P, _I]),
    "orc_btree_free": (None, [_P]),
    "orc_harmonia_build": (_P, [_P, _U64, _I, _U32]),
    "orc_harmonia_build_mode": (_P, [_P, _U64, _I, _U32, _I]),
    "orc_harmonia_depth": (_I, [_P]),
    "orc_harmonia_inner_size": (_U64, [_P]),
    "orc_harmonia_node_count": (_U64, [_P]),
    "orc_harmonia_dump": (None, [_P, _P, _P]),
    "orc_harmonia_lookup": (None, [_P, _P, _U64, _P, _I]),
    "orc_harmonia_free": (None, [_P]),
    "orc_rs_build": (_P, [_P, _U64, _I, _U32, _U32]),
    "orc_rs_num_points": (_U64, [_P]),
    "orc_rs_radix_size": (_U64, [_P]),
    "orc_rs_shift_bits": (_U32, [_P]),
    "orc_rs_dump": (None, [_P, _P, _P, _P]),
    "orc_rs_estimate": (_D, [_P, _U64]),
    "orc_rs_lookup": (None, [_P, _P, _P, _U64, _P, _I]),
    "orc_rs_max_abs_error": (_D, [_P, _P]),
    "orc_rs_free": (None, [_P]),
    "orc_murmur3_64": (_U64, [_U64]),
    "orc_murmur3_32": (_U32, [_U32]),
    "orc_hash_capacity": (_U64, [_U64]),
    "orc_hash_join": (None, [_P, _U64, _I, _P, _U64, _P]),
    "orc_partition": (None, [_P, _U64, _I, _U32, _U32, _P, _P, _P]),
    "orc_join_aggregate": (None, [_P, _P, _U64, _P, _P, _P, _P, _U64, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(_U64), _I]),
    "orc_promo_revenue": (C.c_int64, [C.c_int64, C.c_int64]),
    "orc_zipf_mt19937": (None, [_U64, _D, _U32, _U64, _P]),
    "orc_zipf_trial": (_U64, [_U64, _D, _D]),
    "orc_splitmix64": (_U64, [_U64, _U64]),
    "orc_mlp_forward": (None, [_P, _U32, _P, _U64, _P]),
    "orc_mlp_loss_grad": (_D, [_P, _U32, _P, _P, _U64, _P]),
}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p()


def lower_bound(arr: np.ndarray, key: int) -> int:
    return int(load().orc_lower_bound(_p(arr), arr.size, arr.dtype.itemsize, int(key)))


def branchy_binary_search(arr: np.ndarray, size: int, key: int) -> int:
    return int(load().orc_branchy_binary_search(_p(arr), size, arr.dtype.itemsize, int(key)))


def lookup_lower_bound(col: np.ndarray, probes: np.ndarray, threads: int = 0) -> np.ndarray:
    out = np.empty(probes.size, dtype=np.uint32)
    load().orc_lookup_lower_bound(_p(col), col.size, col.dtype.itemsize, _p(probes), probes.size, _p(out), threads)
    return out


def checksum(rowids: np.ndarray) -> int:
    return int(load().orc_checksum(_p(rowids), rowids.size))


def btree_capacities(key_bytes: int, value_bytes: int):
    a, b = _U32(), _U32()
    load().orc_btree_capacities(key_bytes, value_bytes, C.byref(a), C.byref(b))
    return a.value, b.value


class BTree:
    """The reference's clustered B+-tree (4 KiB pages) over a sorted column."""

    def __init__(self, col: np.ndarray, value_bytes: int = 4):
        self.col = col
        self.h = load().orc_btree_build(_p(col), col.size, col.dtype.itemsize, value_bytes)

    def info(self):
        a, b, d, p = _U32(), _U32(), _U32(), _U64()
        load().orc_btree_info(self.h, C.byref(a), C.byref(b), C.byref(d), C.byref(p))
        return {"leaf_max": a.value, "inner_max": b.value, "depth": d.value, "pages": p.value}

    def lookup(self, probes: np.ndarray, threads: int = 0) -> np.ndarray:
        out = np.empty(probes.size, dtype=np.uint32)
        load().orc_btree_lookup(self.h, _p(probes), probes.size, _p(out), threads)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            load().orc_btree_free(self.h)
            self.h = None


class Harmonia:
    """The reference's Harmonia tree (breadth-first keys + prefix-sum children)."""

    def __init__(self, col: np.ndarray, fanout: int = 33, reuse_input: bool = True):
        self.col = col
        self.h = load().orc_harmonia_build_mode(_p(col), col.size, col.dtype.itemsize, fanout, int(reuse_input))
        if not self.h:
            raise ValueError("harmonia: the dataset size must be a multiple of the node size")

    def depth(self) -> int:
        return int(load().orc_harmonia_depth(self.h))

    def arrays(self):
        lib = load()
        keys = np.empty(lib.orc_harmonia_inner_size(self.h), dtype=np.uint64)
        children = np.empty(lib.orc_harmonia_node_count(self.h), dtype=np.uint64)
        lib.orc_harmonia_dump(self.h, _p(keys), _p(children))
        return keys, children

    def lookup(self, probes: np.ndarray, threads: int = 0) -> np.ndarray:
        out = np.empty(probes.size, dtype=np.uint32)
        load().orc_harmonia_lookup(self.h, _p(probes), probes.size, _p(out), threads)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            load().orc_harmonia_free(self.h)
            self.h = None


class RadixSpline:
    def __init__(self, col: np.ndarray, radix_bits: int = 18, max_error: int = 32):
        self.col = col
        self.h = load().orc_rs_build(_p(col), col.size, col.dtype.itemsize, radix_bits, max_error)

    def arrays(self):
        lib = load()
        n = lib.orc_rs_num_points(self.h)
        r = lib.orc_rs_radix_size(self.h)
        px = np.empty(n, dtype=np.uint64)
        py = np.empty(n, dtype=np.float64)
        radix = np.empty(r, dtype=np.uint32)
        lib.orc_rs_dump(self.h, _p(px), _p(py), _p(radix))
        return px, py, radix

    def shift_bits(self) -> int:
        return int(load().orc_rs_shift_bits(self.h))

    def estimate(self, key: int) -> float:
        return float(load().orc_rs_estimate(self.h, int(key)))

    def max_abs_error(self) -> float:
        return float(load().orc_rs_max_abs_error(self.h, _p(self.col)))

    def lookup(self, probes: np.ndarray, threads: int = 0) -> np.ndarray:
        out = np.empty(probes.size, dtype=np.uint32)
        load().orc_rs_lookup(self.h, _p(self.col), _p(probes), probes.size, _p(out), threads)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            load().orc_rs_free(self.h)
            self.h = None


def murmur3_64(k: int) -> int:
    return int(load().orc_murmur3_64(k))


def murmur3_32(k: int) -> int:
    return int(load().orc_murmur3_32(k))


def hash_capacity(n: int) -> int:
    return int(load().orc_hash_capacity(n))


def hash_join(col: np.ndarray, probes: np.ndarray) -> np.ndarray:
    out = np.empty(probes.size, dtype=np.uint32)
    load().orc_hash_join(_p(col), col.size, col.dtype.itemsize, _p(probes), probes.size, _p(out))
    return out


def partition(keys: np.ndarray, radix_bits: int, ignore_bits: int):
    counts = np.empty(1 << radix_bits, dtype=np.uint64)
    ok = np.empty(keys.size, dtype=keys.dtype)
    op = np.empty(keys.size, dtype=np.uint32)
    load().orc_partition(_p(keys), keys.size, keys.dtype.itemsize, radix_bits, ignore_bits, _p(counts), _p(ok), _p(op))
    return counts, ok, op


def join_aggregate(dim_sorted_key, dim_flag, fkey, date, price, disc, date_lo, date_hi, threads: int = 0):
    num, den, mt = C.c_int64(), C.c_int64(), _U64()
    load().orc_join_aggregate(_p(dim_sorted_key), _p(dim_flag), dim_sorted_key.size, _p(fkey), _p(date), _p(price),
                              _p(disc), fkey.size, date_lo, date_hi, C.byref(num), C.byref(den), C.byref(mt), threads)
    return num.value, den.value, mt.value


def promo_revenue(num: int, den: int) -> int:
    return int(load().orc_promo_revenue(num, den))


def zipf_mt19937(n: int, q: float, seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    load().orc_zipf_mt19937(n, q, seed, count, _p(out))
    return out


def zipf_trial(n: int, q: float, u01: float) -> int:
    return int(load().orc_zipf_trial(n, q, u01))


def splitmix64(seed: int, i: int) -> int:
    return int(load().orc_splitmix64(seed, i))


def mlp_forward(params: np.ndarray, hidden: int, x: np.ndarray) -> np.ndarray:
    out = np.empty(x.size, dtype=np.float32)
    load().orc_mlp_forward(_p(params), hidden, _p(x), x.size, _p(out))
    return out


def mlp_loss_grad(params: np.ndarray, hidden: int, x: np.ndarray, y: np.ndarray):
    grad = np.empty(3 * hidden + 1, dtype=np.float64)
    loss = load().orc_mlp_loss_grad(_p(params), hidden, _p(x), _p(y), x.size, _p(grad))
    return float(loss), grad


def ref_zipf_available() -> bool:
    return os.path.exists(REF_ZIPF_PATH)


def ref_zipf_mt19937(n: int, q: float, seed: int, count: int) -> np.ndarray:
    """The reference's own zipf_distribution (src/zipf.hpp) compiled into _ref/."""
    lib = C.CDLL(REF_ZIPF_PATH)
    lib.ref_zipf_mt19937.restype = None
    lib.ref_zipf_mt19937.argtypes = [_U64, _D, _U32, _U64, _P]
    out = np.empty(count, dtype=np.uint64)
    lib.ref_zipf_mt19937(n, q, seed, count, _p(out))
    return out
