"""gpujoin_amd - MI355X-native index-join probe library (Python host side).

Thin, typed wrappers over the C ABI of libgj.so (include/gj.h).  The Python
layer owns no compute: every lookup, partition, aggregate and training step is
a HIP kernel in libgj.so, and nothing here falls back to the CPU.

Reference interface mirrored (src/indexes.cuh:30-39, src/index_lookup.cuh:11-35):
    Index.build(kind, column)          abstract_index::construct(h_column, d_column)
    Index.memory_consumption()         abstract_index::memory_consumption()
    Index.lookup(probes, approach=...) plain / bws / partitioning / hj approach run
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (APPROACH_NAMES, BATCH_GRAPH, BATCH_LAUNCH, BATCH_PERSISTENT, INDEX_NAMES, NOT_FOUND, AggParams,
                   AggResult, FactBatch, GjError, IndexParams, IndexStats, LookupParams, Partitioned,
                   PartitionParams, ProbeBatch, check)

__all__ = [
    "DeviceArray", "Stream", "Event", "Index", "NOT_FOUND", "GjError", "gen_build_keys", "gen_probes",
    "gen_dimension", "gen_fact", "sort_dimension", "device_info", "gather", "pointer_chase", "stream_copy", "stream_read",
    "index_params", "lookup_params", "partition", "join_aggregate", "promo_revenue", "LookupPlan",
    "AggregatePlan", "mlp_loss_grad", "mlp_forward", "autotune", "autotune_index",
]


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Stream:
    """A non-blocking HIP stream (gj_stream_create)."""

    def __init__(self):
        h = C.c_void_p()
        check(_lib.call("gj_stream_create", C.byref(h)))
        self.handle = h

    def synchronize(self):
        check(_lib.call("gj_stream_synchronize", self.handle))

    def close(self):
        if self.handle:
            _lib.call("gj_stream_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _s(stream: Optional[Stream]) -> C.c_void_p:
    return stream.handle if stream is not None else C.c_void_p()


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(_lib.call("gj_event_create", C.byref(h)))
        self.handle = h

    def record(self, stream: Optional[Stream] = None):
        check(_lib.call("gj_event_record", self.handle, _s(stream)))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        check(_lib.call("gj_event_elapsed_ms", self.handle, later.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.call("gj_event_destroy", self.handle)
        except Exception:
            pass


class DeviceArray:
    """Owning HBM buffer of `count` elements of `dtype` (the role of device_array_wrapper)."""

    def __init__(self, count: int, dtype, stream: Optional[Stream] = None):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        check(_lib.call("gj_malloc", C.byref(p), max(self.nbytes, 16)))
        self.ptr = p
        self.stream = stream

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream: Optional[Stream] = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, stream)
        if a.nbytes:
            check(_lib.call("gj_memcpy_htod", d.ptr, _ptr(a), a.nbytes, _s(stream)))
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        if a.nbytes:
            check(_lib.call("gj_memcpy_htod", self.ptr, _ptr(a), a.nbytes, _s(self.stream)))

    def numpy(self, count: Optional[int] = None) -> np.ndarray:
        n = self.count if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if out.nbytes:
            check(_lib.call("gj_memcpy_dtoh", _ptr(out), self.ptr, out.nbytes, _s(self.stream)))
        return out

    def fill(self, byte: int, stream: Optional[Stream] = None):
        check(_lib.call("gj_memset", self.ptr, int(byte), max(self.nbytes, 16), _s(stream)))

    @property
    def addr(self) -> int:
        return int(self.ptr.value or 0)

    def offset(self, elements: int) -> C.c_void_p:
        return C.c_void_p(self.addr + int(elements) * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            _lib.call("gj_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostArray:
    """Page-locked host memory (gj_host_malloc), mapped into the GPU's address space, so
    kernels read it in place over the host link: an index column that stays in host
    memory (the reference's index_lookup_config placement, src/index_lookup_config.hpp:52-66).
    Same attributes as DeviceArray (ptr, dtype, count) where an index takes its column."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = C.c_void_p()
        check(_lib.call("gj_host_malloc", C.byref(p), max(self.nbytes, 16)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "HostArray":
        a = np.ascontiguousarray(a)
        h = cls(a.size, a.dtype)
        h.view()[:] = a
        return h

    def view(self) -> np.ndarray:
        buf = (C.c_uint8 * max(self.nbytes, 1)).from_address(int(self.ptr.value))
        return np.frombuffer(buf, dtype=np.uint8, count=self.nbytes).view(self.dtype)

    def free(self):
        if self.ptr:
            _lib.call("gj_host_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cu = C.c_int()
    major = C.c_int()
    check(_lib.call("gj_device_info", name, 256, C.byref(cu), C.byref(major)))
    return {"name": name.value.decode(), "compute_units": cu.value, "major": major.value}


def device_count() -> int:
    c = C.c_int()
    check(_lib.call("gj_device_count", C.byref(c)))
    return c.value


def synchronize():
    check(_lib.call("gj_device_synchronize"))


def index_params(**kw) -> IndexParams:
    p = IndexParams()
    _lib.call("gj_index_params_default", C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def lookup_params(approach="plain", **kw) -> LookupParams:
    p = LookupParams()
    _lib.call("gj_lookup_params_default", C.byref(p))
    p.approach = APPROACH_NAMES[approach] if isinstance(approach, str) else int(approach)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Index:
    """A built index over a sorted device column (gj_index_build).

    The column is borrowed: keep `column` alive while the index exists.
    """

    def __init__(self, kind, column: DeviceArray, host_column: Optional[np.ndarray] = None,
                 params: Optional[IndexParams] = None, stream: Optional[Stream] = None, n: Optional[int] = None):
        self.kind = INDEX_NAMES[kind] if isinstance(kind, str) else int(kind)
        self.column = column
        self.key_bytes = column.dtype.itemsize
        self.n = column.count if n is None else int(n)
        h = C.c_void_p()
        hp = _ptr(host_column) if host_column is not None else C.c_void_p()
        p = C.byref(params) if params is not None else None
        check(_lib.call("gj_index_build", self.kind, hp, column.ptr, self.n, self.key_bytes, p, _s(stream),
                        C.byref(h)))
        self.handle = h

    def lookup(self, probes: DeviceArray, out: DeviceArray, m: Optional[int] = None,
               params: Optional[LookupParams] = None, stream: Optional[Stream] = None, approach=None, **kw):
        if params is None:
            params = lookup_params(approach or "plain", **kw)
        mm = probes.count if m is None else int(m)
        check(_lib.call("gj_index_lookup", self.handle, probes.ptr, mm, out.ptr, C.byref(params), _s(stream)))

    def lookup_ptr(self, probes_ptr, m: int, out_ptr, params: LookupParams, stream: Optional[Stream] = None):
        check(_lib.call("gj_index_lookup", self.handle, probes_ptr, int(m), out_ptr, C.byref(params), _s(stream)))

    def memory_consumption(self) -> int:
        return int(_lib.call("gj_index_memory", self.handle))

    def stats(self) -> dict:
        st = IndexStats()
        check(_lib.call("gj_index_get_stats", self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in IndexStats._fields_}

    def close(self):
        if self.handle:
            _lib.call("gj_index_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- generators (host, seeded) -------------------------------------------

def _kdtype(key_bytes: int):
    return np.uint64 if key_bytes == 8 else np.uint32


def gen_build_keys(n: int, key_bytes: int = 8, dist: int = _lib.KEYS_UNIFORM, upper_bits: int = 0, seed: int = 1,
                   threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=_kdtype(key_bytes))
    check(_lib.call("gj_gen_build_keys", _ptr(out), n, key_bytes, dist, upper_bits, seed, threads))
    return out


def gen_probes(keys: np.ndarray, m: int, dist: int = _lib.PROBES_UNIFORM, zipf_theta: float = 1.0,
               miss_fraction: float = 0.0, upper_bits: int = 0, seed: int = 2, threads: int = 0,
               expected: bool = True):
    kb = keys.dtype.itemsize
    out = np.empty(m, dtype=keys.dtype)
    exp = np.empty(m, dtype=np.uint32) if expected else None
    check(_lib.call("gj_gen_probes", _ptr(keys), keys.size, kb, _ptr(out), m, dist, zipf_theta, miss_fraction,
                    upper_bits, seed, _ptr(exp) if exp is not None else None, threads))
    return (out, exp) if expected else out


def gen_dimension(rows: int, seed: int = 0x41, threads: int = 0):
    key = np.empty(rows, dtype=np.int32)
    flag = np.empty(rows, dtype=np.uint8)
    check(_lib.call("gj_gen_dimension", _ptr(key), _ptr(flag), rows, seed, threads))
    return key, flag


def gen_fact(rows: int, dim_rows: int, days: int = 2557, price_lo: int = 90000, price_hi: int = 10500000,
             disc_max: int = 10, seed: int = 0x42, threads: int = 0):
    fkey = np.empty(rows, dtype=np.int32)
    date = np.empty(rows, dtype=np.int32)
    price = np.empty(rows, dtype=np.int64)
    disc = np.empty(rows, dtype=np.int8)
    check(_lib.call("gj_gen_fact", _ptr(fkey), _ptr(date), _ptr(price), _ptr(disc), rows, dim_rows, days, price_lo,
                    price_hi, disc_max, seed, threads))
    return fkey, date, price, disc


def sort_dimension(key: np.ndarray, flag: np.ndarray):
    sk = np.empty(key.size, dtype=np.uint32)
    sf = np.empty(key.size, dtype=np.uint8)
    check(_lib.call("gj_sort_dimension", _ptr(key), _ptr(flag), key.size, _ptr(sk), _ptr(sf)))
    return sk, sf


def rng_u64(seed: int, index: int) -> int:
    return int(_lib.call("gj_rng_u64", seed, index))


# ---- microbenchmarks -------------------------------------------------------

def gather(array: DeviceArray, probes: DeviceArray, out: DeviceArray, m: Optional[int] = None, block: int = 256,
           items: int = 4, stream: Optional[Stream] = None):
    mm = probes.count if m is None else int(m)
    check(_lib.call("gj_gather", array.ptr, array.count, probes.ptr, mm, out.ptr, block, items, _s(stream)))


def pointer_chase(nxt: DeviceArray, chains: int, steps: int, out: DeviceArray, stream: Optional[Stream] = None,
                  start: Optional[DeviceArray] = None):
    """cur = nxt[cur], `steps` times per chain; chain c starts at start[c] (start may be `out`
    itself, to continue each chain where the last call stopped) or at a spread default."""
    sp = start.ptr if start is not None else C.c_void_p()
    check(_lib.call("gj_pointer_chase", nxt.ptr, nxt.count, chains, steps, sp, out.ptr, _s(stream)))


def stream_copy(src: DeviceArray, dst: DeviceArray, nbytes: int, stream: Optional[Stream] = None):
    check(_lib.call("gj_stream_copy", src.ptr, dst.ptr, nbytes, _s(stream)))


def stream_read(src: DeviceArray, nbytes: int, out: DeviceArray, stream: Optional[Stream] = None):
    """XORs every 8-byte word of the first nbytes of src into out[0] (a uint64 the caller sets)."""
    check(_lib.call("gj_stream_read", src.ptr, nbytes, out.ptr, _s(stream)))


# ---- partitioner / join-aggregate / plans / learned hooks ------------------

def partition(keys: DeviceArray, m: int, radix_bits: int, ignore_bits: int, passes: int = 1,
              stream: Optional[Stream] = None, dropped: Optional[DeviceArray] = None):
    """Radix-partitions m keys; returns (keys, payloads, offsets, counts) as device arrays.
    `dropped` (one uint64, optional) counts tuples that did not fit the output: 0 unless a fault."""
    kb = keys.dtype.itemsize
    pp = PartitionParams(radix_bits, ignore_bits, passes)
    cap = int(_lib.call("gj_partition_capacity", m, kb, radix_bits))
    scratch_bytes = int(_lib.call("gj_partition_scratch_bytes", m, kb, C.byref(pp)))
    dk = DeviceArray(cap, keys.dtype)
    dp = DeviceArray(cap, np.uint32)
    off = DeviceArray(1 << radix_bits, np.uint64)
    cnt = DeviceArray(1 << radix_bits, np.uint64)
    scratch = DeviceArray(max(scratch_bytes, 16), np.uint8)
    out = Partitioned(dk.ptr, dp.ptr, off.ptr, cnt.ptr, cap, dropped.ptr if dropped is not None else None)
    check(_lib.call("gj_partition", keys.ptr, m, kb, C.byref(pp), C.byref(out), scratch.ptr, scratch_bytes,
                    _s(stream)))
    return dk, dp, off, cnt


def join_aggregate(index: Index, fkey: DeviceArray, date: DeviceArray, price: DeviceArray, disc: DeviceArray,
                   flag: DeviceArray, date_lo: int, date_hi: int, result: DeviceArray, rows: Optional[int] = None,
                   offset: int = 0, stream: Optional[Stream] = None, strategy="lane_refill", radix_bits: int = 0):
    """Adds the Q14-shaped sums of the fact rows to `result`; strategy is "lane_refill" (default),
    "plain", "partitioned", "sorted" or "lookup_first" (gj_agg_strategy)."""
    r = fkey.count if rows is None else int(rows)
    fb = FactBatch(fkey.offset(offset), date.offset(offset), price.offset(offset), disc.offset(offset), r)
    st = _lib.AGG_STRATEGIES[strategy] if isinstance(strategy, str) else int(strategy)
    ap = AggParams(date_lo, date_hi, flag.ptr, st, radix_bits)
    check(_lib.call("gj_join_aggregate", index.handle, C.byref(fb), C.byref(ap), result.ptr, _s(stream)))


def promo_revenue(num: int, den: int) -> int:
    return int(_lib.call("gj_promo_revenue", num, den))


class _Plan:
    def run(self, stream: Optional[Stream] = None):
        check(_lib.call("gj_batch_plan_run", self.handle, _s(stream)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.call("gj_batch_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LookupPlan(_Plan):
    """Uneven probe batches against one resident index (gj_lookup_plan_create)."""

    def __init__(self, index: Index, batches, mode: int, stream: Optional[Stream] = None):
        arr = (ProbeBatch * len(batches))(*[ProbeBatch(p, o, m) for (p, o, m) in batches])
        h = C.c_void_p()
        check(_lib.call("gj_lookup_plan_create", index.handle, arr, len(batches), mode, _s(stream), C.byref(h)))
        self.handle = h
        self.index = index


class AggregatePlan(_Plan):
    def __init__(self, index: Index, batches, flag: DeviceArray, date_lo: int, date_hi: int, results: DeviceArray,
                 mode: int, stream: Optional[Stream] = None):
        arr = (FactBatch * len(batches))(*[FactBatch(*b) for b in batches])
        ap = AggParams(date_lo, date_hi, flag.ptr, _lib.AGG_LANE_REFILL, 0)
        h = C.c_void_p()
        check(_lib.call("gj_aggregate_plan_create", index.handle, arr, len(batches), C.byref(ap), results.ptr, mode,
                        _s(stream), C.byref(h)))
        self.handle = h
        self.index = index


def autotune(index: Index, probes: DeviceArray, scratch: DeviceArray, m: Optional[int] = None,
             stream: Optional[Stream] = None):
    best = lookup_params()
    rate = C.c_double()
    mm = probes.count if m is None else int(m)
    check(_lib.call("gj_autotune_lookup", index.handle, probes.ptr, mm, scratch.ptr, C.byref(best), C.byref(rate),
                    _s(stream)))
    return best, rate.value


def autotune_index(kind, column: DeviceArray, probes: DeviceArray, scratch: DeviceArray,
                   host_column: Optional[np.ndarray] = None, base: Optional[IndexParams] = None,
                   m: Optional[int] = None, n: Optional[int] = None, stream: Optional[Stream] = None):
    """Index parameters (threads per lookup, LDS levels, radix bits, error bounds, leaves) of
    the fastest candidate build of `kind` over `column` (gj_autotune_index); returns (params, rate)."""
    k = INDEX_NAMES[kind] if isinstance(kind, str) else int(kind)
    best = IndexParams()
    rate = C.c_double()
    nn = column.count if n is None else int(n)
    mm = probes.count if m is None else int(m)
    hp = _ptr(host_column) if host_column is not None else C.c_void_p()
    bp = C.byref(base) if base is not None else None
    check(_lib.call("gj_autotune_index", k, hp, column.ptr, nn, column.dtype.itemsize, bp, probes.ptr, mm,
                    scratch.ptr, C.byref(best), C.byref(rate), _s(stream)))
    return best, rate.value


def mlp_loss_grad(params: DeviceArray, hidden: int, x: DeviceArray, y: DeviceArray, grad: DeviceArray,
                  loss: DeviceArray, n: Optional[int] = None, stream: Optional[Stream] = None):
    nn = x.count if n is None else int(n)
    check(_lib.call("gj_mlp_loss_grad", params.ptr, hidden, x.ptr, y.ptr, nn, grad.ptr, loss.ptr, _s(stream)))


def mlp_forward(params: DeviceArray, hidden: int, x: DeviceArray, out: DeviceArray, n: Optional[int] = None,
                stream: Optional[Stream] = None):
    nn = x.count if n is None else int(n)
    check(_lib.call("gj_mlp_forward", params.ptr, hidden, x.ptr, nn, out.ptr, _s(stream)))
