"""Measurement suite over the workload configurations beyond the headline.

Run with `python bench.py --suite [--quick]`.  Every measurement checks its row ids
or sums against the generator's ground truth (or the oracle for the aggregates)
and records lookups/s (rows/s for the aggregate) next to the CPU baseline.
Writes one JSON object per measurement to stdout-compatible lines via `emit`.

  u32 probe      2^20 unique u32 keys from [0, 2^32), 2^24 uniform probes: row ids + checksum
  Zipf probe     2^26 u64 keys, 2^28 probes, Zipf theta = 1.0 (rank 0 never drawn, as in the reference)
  sorted / dense probes in key order; dense keys (key = row id) as in the reference's scripts
  lookup plans   uneven probe batches through launches, a hipGraph and the persistent kernel
  host column    the index column in page-locked host memory, plain vs partitioned
  partitioned    fan-out 2^8 / 2^10 / 2^12, 1 and 2 passes, windows 2^18..2^26;
                 index-size sweep 2^20..2^30 keys (B-tree and learned)
  learned        RadixSpline err 8/32/64 and the PLR / MLP models over uniform,
                 lognormal and clustered keys, 10 % misses; B-tree fallbacks counted
  join-aggregate 2e6-row dimension, 6e7 fact rows, 30-day window,
                 log-uniform batches of 1e3..1e7 rows, per-batch launches / hipGraph / persistent kernel
"""
from __future__ import annotations

import json
import os
import statistics
import sys
import time

import numpy as np

from gpujoin_amd import _lib


def _time(gj, fn, stream, steps=5, warmup=2):
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = gj.Event(), gj.Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / steps


def run(gj, G, quick=False, cpu=True, emit=None):
    import oracle

    emit = emit or (lambda d: print(json.dumps(d), flush=True))
    stream = gj.Stream()
    cores = os.cpu_count() or 1
    log = lambda *a: print("[suite]", *a, file=sys.stderr, flush=True)

    # ---- u32 probe -------------------------------------------------------------------
    keys = gj.gen_build_keys(1 << 20, 4, _lib.KEYS_UNIFORM, 0, 0x1)
    probes, exp = gj.gen_probes(keys, 1 << 24, _lib.PROBES_UNIFORM, 1.0, 0.0, 0, 0x2)
    dk, dp = gj.DeviceArray.from_numpy(keys), gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for kind in ("binary_search", "btree", "learned_plr"):
        idx = gj.Index(kind, dk, host_column=keys, stream=stream)
        ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params(), stream=stream), stream)
        got = out.numpy()
        rec = {"config": "u32 probe: 2^20 u32 keys, 2^24 uniform probes", "index": kind, "lookups_per_s": probes.size / ms * 1e3,
               "ms": ms, "rowids_match": bool(np.array_equal(got, exp)),
               "checksum": int(got.astype(np.uint64).sum()), "checksum_expected": int(exp.astype(np.uint64).sum())}
        if cpu and kind == "binary_search":
            t = time.perf_counter()
            ref = oracle.lookup_lower_bound(keys, probes, threads=cores)
            rec["cpu_lookups_per_s"] = probes.size / (time.perf_counter() - t)
            rec["cpu_cores"] = cores
            rec["cpu_checksum"] = oracle.checksum(ref)
        emit(rec)
        idx.close()
    del dk, dp, out

    n_log = 24 if quick else 26
    m_log = 26 if quick else 28
    n, m = 1 << n_log, 1 << m_log
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x11)
    dk = gj.DeviceArray.from_numpy(keys, stream)
    out = gj.DeviceArray(m, np.uint32)

    # ---- Zipf probe --------------------------------------------------------------
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_ZIPF, 1.0, 0.0, 63, 0x13)
    dp = gj.DeviceArray.from_numpy(probes, stream)
    for kind in ("binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "hash"):
        idx = gj.Index(kind, dk, host_column=keys, stream=stream)
        for appr in ("plain", "partitioning"):
            ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params(appr), stream=stream), stream, steps=3)
            emit({"config": f"Zipf 1.0 probe: 2^{n_log} u64 keys, 2^{m_log} probes", "index": kind, "approach": appr,
                  "lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G, "ms": ms,
                  "rowids_match": bool(np.array_equal(out.numpy(), exp))})
        idx.close()
    del probes, exp

    # ---- hash-join driver: a linear-probing table built over the column on every call ----
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x14)
    dp.upload(probes)
    idx = gj.Index("btree", dk, host_column=keys, stream=stream)
    ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params("hj"), stream=stream), stream, steps=3, warmup=1)
    emit({"config": f"hash-join driver (table build + probe per call): 2^{n_log} u64 keys, 2^{m_log} uniform probes",
          "index": "btree", "approach": "hj", "lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G,
          "ms": ms, "rowids_match": bool(np.array_equal(out.numpy(), exp))})
    idx.close()
    del probes, exp

    # ---- sorted probes (the reference's -s true) and dense keys (its --dataset dense) ----
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x17)
    order = np.argsort(probes, kind="stable")
    dp.upload(probes[order])
    exp_sorted = exp[order]
    del probes, exp, order
    for kind in ("btree", "learned_plr", "radix_spline", "harmonia"):
        idx = gj.Index(kind, dk, host_column=keys, stream=stream)
        ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params("plain"), stream=stream), stream, steps=3)
        emit({"config": f"sorted probes: 2^{n_log} u64 keys, 2^{m_log} uniform probes in key order", "index": kind,
              "lookups_per_s": m / ms * 1e3, "ms": ms, "rowids_match": bool(np.array_equal(out.numpy(), exp_sorted))})
        idx.close()
    del exp_sorted
    dense = gj.gen_build_keys(n, 8, _lib.KEYS_DENSE, 0, 0)
    dd = gj.DeviceArray.from_numpy(dense, stream)
    probes, exp = gj.gen_probes(dense, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 0, 0x18)
    dp.upload(probes)
    for kind in ("binary_search", "btree", "harmonia", "radix_spline", "learned_plr"):
        idx = gj.Index(kind, dd, host_column=dense, stream=stream)
        ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params("plain"), stream=stream), stream, steps=3)
        emit({"config": f"dense keys (key = row id): 2^{n_log} u64 keys, 2^{m_log} uniform probes", "index": kind,
              "lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G, "ms": ms,
              "rowids_match": bool(np.array_equal(out.numpy(), exp))})
        idx.close()
    del dd, dense, probes, exp

    # ---- lookup plans: uneven probe batches against one resident index ----------------
    rng = np.random.default_rng(0x15)
    sizes, total = [], 0
    while total < m:
        sz = min(int(np.exp(rng.uniform(np.log(1e3), np.log(1e7)))), m - total)
        sizes.append(sz)
        total += sz
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x16)
    dp.upload(probes)
    idx = gj.Index("learned_plr", dk, host_column=keys, stream=stream)
    batches, off = [], 0
    for sz in sizes:
        batches.append((dp.offset(off), out.offset(off), sz))
        off += sz
    for mode, mname in ((_lib.BATCH_LAUNCH, "launch"), (_lib.BATCH_GRAPH, "hipgraph"),
                        (_lib.BATCH_PERSISTENT, "persistent")):
        plan = gj.LookupPlan(idx, batches, mode, stream=stream)
        ms = _time(gj, lambda: plan.run(stream), stream, steps=3, warmup=1)
        emit({"config": f"lookup plan, {len(sizes)} log-uniform probe batches (1e3..1e7), 2^{n_log} u64 keys",
              "index": "learned_plr", "mode": mname, "lookups_per_s": m / ms * 1e3, "ms": ms,
              "rowids_match": bool(np.array_equal(out.numpy(), exp))})
        plan.close()
    idx.close()
    del probes, exp

    # ---- partitioned sweeps -------------------------------------------------
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x21)
    dp.upload(probes)
    for kind in ("btree", "learned_plr"):
        idx = gj.Index(kind, dk, host_column=keys, stream=stream)
        for bits, passes in ((8, 1), (10, 1), (12, 1), (12, 2)):
            for wlog in ((18, 20, 22, 24, 26) if not quick else (20, 24)):
                lp = gj.lookup_params("partitioning", radix_bits=bits, passes=passes, window=1 << wlog)
                ms = _time(gj, lambda: idx.lookup(dp, out, params=lp, stream=stream), stream, steps=3, warmup=1)
                emit({"config": f"partitioned probe: 2^{n_log} u64 keys, 2^{m_log} uniform probes", "index": kind,
                      "fanout": 1 << bits, "passes": passes, "window": 1 << wlog, "lookups_per_s": m / ms * 1e3,
                      "gather_roofline_frac": m / ms * 1e3 / G, "ms": ms,
                      "rowids_match": bool(np.array_equal(out.numpy(), exp))})
        idx.close()
    del dk
    sizes = (20, 22, 24, 26, 28, 30) if not quick else (20, 24)
    for s_log in sizes:
        k2 = gj.gen_build_keys(1 << s_log, 8, _lib.KEYS_UNIFORM, 63, 0x22 + s_log)
        p2, e2 = gj.gen_probes(k2, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x23 + s_log)
        dk2 = gj.DeviceArray.from_numpy(k2, stream)
        dp.upload(p2)
        for kind in ("btree", "learned_plr"):
            idx = gj.Index(kind, dk2, host_column=k2, stream=stream)
            for appr in ("plain", "partitioning"):
                ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params(appr), stream=stream), stream,
                           steps=3, warmup=1)
                emit({"config": f"index-size sweep: 2^{s_log} u64 keys ({(8 << s_log) / 2**30:.2f} GiB), "
                                f"2^{m_log} probes", "index": kind, "approach": appr, "lookups_per_s": m / ms * 1e3,
                      "ms": ms, "rowids_match": bool(np.array_equal(out.numpy(), e2))})
            idx.close()
        del dk2, k2, p2, e2
    log("partitioned sweeps done")

    # ---- index column in host memory (page-locked, read in place over the host link) ----
    mh_log = 22 if quick else 24
    mh = 1 << mh_log
    kh = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x24)
    ph, eh = gj.gen_probes(kh, mh, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x25)
    hcol = gj.HostArray.from_numpy(kh)
    dph = gj.DeviceArray.from_numpy(ph, stream)
    doh = gj.DeviceArray(mh, np.uint32)
    for kind in ("btree", "learned_plr"):
        idx = gj.Index(kind, hcol, host_column=kh, stream=stream)
        for appr, kw in (("plain", {}), ("partitioning", {"radix_bits": 12, "window": mh}),
                         ("partitioning", {"radix_bits": 12, "passes": 2, "window": mh})):
            lp = gj.lookup_params(appr, **kw)
            ms = _time(gj, lambda: idx.lookup(dph, doh, params=lp, stream=stream), stream, steps=2, warmup=1)
            emit({"config": f"column in host memory: 2^{n_log} u64 keys in page-locked host memory, "
                            f"2^{mh_log} uniform probes in HBM", "index": kind, "approach": appr,
                  "passes": kw.get("passes", 1), "fanout": 1 << kw.get("radix_bits", 0) if kw else None,
                  "lookups_per_s": mh / ms * 1e3, "ms": ms, "rowids_match": bool(np.array_equal(doh.numpy(), eh))})
        idx.close()
    hcol.free()
    del kh, ph, eh, dph, doh
    log("host-memory column done")

    # ---- learned indexes with misses ------------------------------------------------
    for dist, dname in ((_lib.KEYS_UNIFORM, "uniform"), (_lib.KEYS_LOGNORMAL, "lognormal"),
                        (_lib.KEYS_CLUSTERED, "clustered")):
        k4 = gj.gen_build_keys(n, 8, dist, 63, 0x31 + dist)
        p4, e4 = gj.gen_probes(k4, m, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x32 + dist)
        dk4 = gj.DeviceArray.from_numpy(k4, stream)
        dp.upload(p4)
        variants = [("btree", {}), ("radix_spline", {"rs_max_error": 8}), ("radix_spline", {"rs_max_error": 32}),
                    ("radix_spline", {"rs_max_error": 64}), ("learned_plr", {}), ("learned_mlp", {})]
        for kind, prm in variants:
            t = time.time()
            idx = gj.Index(kind, dk4, host_column=k4, params=gj.index_params(**prm), stream=stream)
            build_s = time.time() - t
            st = idx.stats()
            ms = _time(gj, lambda: idx.lookup(dp, out, params=gj.lookup_params(), stream=stream), stream, steps=3)
            emit({"config": f"learned, 10% misses: 2^{n_log} {dname} u64 keys, 2^{m_log} probes, 10% misses", "index": kind,
                  "params": prm, "lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G, "ms": ms,
                  "rowids_match": bool(np.array_equal(out.numpy(), e4)), "build_s": build_s,
                  "fallback_keys": st["fallback_keys"], "fallback_leaves": st["fallback_leaves"],
                  "avg_window": st["avg_window"], "max_window": st["max_window"],
                  "spline_points": st["spline_points"], "train_loss": st["train_loss"]})
            idx.close()
        del dk4, k4, p4, e4
    log("learned sweeps done")
    del dp, out
    run_join_aggregate(gj, quick=quick, cpu=cpu, emit=emit, stream=stream)


def run_join_aggregate(gj, quick=False, cpu=True, emit=None, stream=None):
    """The join-aggregate measurements alone: one batch, then log-uniform batches
    through per-batch launches, a replayed hipGraph and the persistent kernel."""
    import oracle

    emit = emit or (lambda d: print(json.dumps(d), flush=True))
    stream = stream or gj.Stream()
    cores = os.cpu_count() or 1
    log = lambda *a: print("[suite]", *a, file=sys.stderr, flush=True)
    dim_rows = 2_000_000
    fact_rows = 6_000_000 if quick else 60_000_000
    key, flag = gj.gen_dimension(dim_rows, 0x41)
    skey, sflag = gj.sort_dimension(key, flag)
    fkey, date, price, disc = gj.gen_fact(fact_rows, dim_rows, 2557, 90000, 10500000, 10, 0x42)
    lo, hi = 1200, 1230
    onum, oden, omt = oracle.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi, threads=cores)
    n_qual = int(np.count_nonzero((date >= lo) & (date < hi)))
    dcols = [gj.DeviceArray.from_numpy(a, stream) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag, stream)
    dkey = gj.DeviceArray.from_numpy(skey, stream)
    rng = np.random.default_rng(0x43)
    sizes, total = [], 0
    while total < fact_rows:
        s = int(np.exp(rng.uniform(np.log(1e3), np.log(1e7))))
        s = min(s, fact_rows - total)
        sizes.append(s)
        total += s
    idx = gj.Index("learned_plr", dkey, host_column=skey, stream=stream)
    res1 = gj.DeviceArray(4, np.int64)

    def single():
        res1.fill(0, stream)
        gj.join_aggregate(idx, *dcols, dflag, lo, hi, res1, stream=stream)

    ms = _time(gj, single, stream)
    got = tuple(int(v) for v in res1.numpy()[:3])
    emit({"config": f"join-aggregate, one batch of {fact_rows} fact rows", "index": "learned_plr",
          "rows_per_s": fact_rows / ms * 1e3, "ms": ms, "sums_match": got == (onum, oden, omt),
          "promo_revenue": gj.promo_revenue(got[0], got[1]),
          # every row's date (4 B), each qualifying row's fkey, price and discount
          # (4 + 8 + 1 B), each match's flag (1 B)
          "algorithmic_GBps": (fact_rows * 4 + n_qual * 13 + omt) / (ms * 1e-3) / 1e9})
    batches, off = [], 0
    for s in sizes:
        batches.append(tuple(c.offset(off) for c in dcols) + (s,))
        off += s
    res = gj.DeviceArray(4 * len(batches), np.int64)
    for mode, mname in ((_lib.BATCH_LAUNCH, "launch"), (_lib.BATCH_GRAPH, "hipgraph"),
                        (_lib.BATCH_PERSISTENT, "persistent")):
        plan = gj.AggregatePlan(idx, batches, dflag, lo, hi, res, mode, stream=stream)
        ms = _time(gj, lambda: plan.run(stream), stream)
        sums = res.numpy().reshape(-1, 4)[:, :3].sum(axis=0)
        emit({"config": f"join-aggregate, {len(sizes)} log-uniform batches (1e3..1e7 rows)", "mode": mname,
              "rows_per_s": fact_rows / ms * 1e3, "ms": ms,
              "sums_match": tuple(int(v) for v in sums) == (onum, oden, omt)})
        plan.close()
    if cpu:
        t = time.perf_counter()
        oracle.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi, threads=cores)
        cpu_s = time.perf_counter() - t
        emit({"config": "join-aggregate CPU baseline (oracle restatement of the Q14 loop, OpenMP)", "cores": cores,
              "rows_per_s": fact_rows / cpu_s})
    log("join-aggregate done")


# ---- the other configurations' legs of the default bench run ------------------------------

def _timed(gj, fn, stream, steps, warmup):
    """Average ms per call over `steps` calls after `warmup`, from events on `stream`."""
    return _time(gj, fn, stream, steps=steps, warmup=warmup)


def _cpu_median(fn, warmup=3, reps=10):
    """The reference's measure() protocol (src/measuring.hpp:12): warm-ups, then the median of reps."""
    out = None
    for _ in range(warmup):
        out = fn()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), out


ALL_KINDS = ("binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash")


def extras(gj, G, stream, keys64, d_keys64, d_out, cpu=True, steps=5, warmup=2):
    """C1, C2(b), C4 and C5 measured in the default run, every result checked."""
    import oracle

    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    cores = os.cpu_count() or 1
    res = {}

    # C1: 2^20 unique u32 keys, 2^24 uniform probes (100 % hit): row ids + u64 checksum
    keys, (probes, exp) = None, (None, None)
    keys = gj.gen_build_keys(1 << 20, 4, _lib.KEYS_UNIFORM, 0, 0x1)
    probes, exp = gj.gen_probes(keys, 1 << 24, _lib.PROBES_UNIFORM, 1.0, 0.0, 0, 0x2)
    m = probes.size
    dk, dp = gj.DeviceArray.from_numpy(keys, stream), gj.DeviceArray.from_numpy(probes, stream)
    want_sum = int(exp.astype(np.uint64).sum())
    c1 = {"workload": "2^20 unique u32 keys from [0, 2^32), 2^24 uniform u32 probes (100% hit), u32 row ids",
          "checksum_expected": want_sum, "indexes": {}}
    for kind in ALL_KINDS:
        idx = gj.Index(kind, dk, host_column=keys, stream=stream)
        ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=gj.lookup_params(), stream=stream), stream, steps,
                    warmup)
        got = d_out.numpy(m)
        c1["indexes"][kind] = {"lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G,
                               "rowids_match": bool(np.array_equal(got, exp)),
                               "checksum": int(got.astype(np.uint64).sum())}
        idx.close()
        log(f"C1 {kind}: {m / ms * 1e-6:.2f} G lookups/s match={c1['indexes'][kind]['rowids_match']}")
    if cpu:
        sec, ref = _cpu_median(lambda: oracle.lookup_lower_bound(keys, probes, threads=cores))
        c1["cpu_baseline"] = {"lookups_per_s": m / sec, "cores": cores, "kind": "port",
                              "what": "equality-checked lower_bound over all 2^24 probes, median of 10 after 3 "
                                      "warm-ups", "checksum": oracle.checksum(ref)}
    res["C1_u32"] = c1
    del dk, dp, keys, probes, exp

    # C2(b): the headline's 2^26 keys, 2^28 Zipf(1.0) probes over key rank
    m = d_out.count
    probes, exp = gj.gen_probes(keys64, m, _lib.PROBES_ZIPF, 1.0, 0.0, 63, 0x13)
    dp = gj.DeviceArray.from_numpy(probes, stream)
    del probes
    c2 = {"workload": "2^26 unique u64 keys, 2^28 probes Zipf theta=1.0 over key rank (rank 0 never drawn, as in "
                      "the reference), 100% hit", "indexes": {}}
    for kind in ALL_KINDS:
        idx = gj.Index(kind, d_keys64, host_column=keys64, stream=stream)
        ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=gj.lookup_params(), stream=stream), stream, 3, 1)
        ok = bool(np.array_equal(d_out.numpy(m), exp))
        c2["indexes"][kind] = {"lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G,
                               "rowids_match": ok}
        idx.close()
        log(f"C2b {kind}: {m / ms * 1e-6:.2f} G lookups/s match={ok}")
    res["C2b_zipf"] = c2
    del dp, exp

    # C4: learned indexes over three key distributions, 10 % misses, B-tree fallback counted
    c4 = {"workload": "2^26 unique u64 keys (uniform, lognormal(0, 2), 64 Gaussian clusters), 2^28 probes with "
                      "10% misses", "distributions": {}}
    # the learned index over the column itself (no key copy, a few MiB of index) beside the
    # default lines layout; RadixSpline with its radix table staged in LDS (2^13 entries)
    variants = [("btree", "btree", {}), ("radix_spline_err8", "radix_spline", {"rs_max_error": 8}),
                ("radix_spline_err32", "radix_spline", {"rs_max_error": 32}),
                ("radix_spline_err64", "radix_spline", {"rs_max_error": 64}),
                ("radix_spline_lds13_err32", "radix_spline", {"rs_radix_bits": 13, "rs_max_error": 32}),
                ("learned_plr", "learned_plr", {}), ("learned_mlp", "learned_mlp", {}),
                ("learned_plr_column", "learned_plr", {"learned_layout": _lib.LAYOUT_COLUMN}),
                ("learned_mlp_column", "learned_mlp", {"learned_layout": _lib.LAYOUT_COLUMN})]
    for dist, dname in ((_lib.KEYS_UNIFORM, "uniform"), (_lib.KEYS_LOGNORMAL, "lognormal"),
                        (_lib.KEYS_CLUSTERED, "clustered")):
        k4 = gj.gen_build_keys(1 << 26, 8, dist, 63, 0x31 + dist)
        p4, e4 = gj.gen_probes(k4, m, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x32 + dist)
        dk4 = gj.DeviceArray.from_numpy(k4, stream)
        dp4 = gj.DeviceArray.from_numpy(p4, stream)
        del p4
        rows = {}
        for label, kind, prm in variants:
            t = time.time()
            idx = gj.Index(kind, dk4, host_column=k4, params=gj.index_params(**prm), stream=stream)
            build_s = time.time() - t
            st = idx.stats()
            ms = _timed(gj, lambda: idx.lookup(dp4, d_out, m=m, params=gj.lookup_params(), stream=stream), stream, 3,
                        1)
            ok = bool(np.array_equal(d_out.numpy(m), e4))
            rows[label] = {"lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G,
                           "rowids_match": ok, "build_s": build_s, "fallback_keys": st["fallback_keys"],
                           "fallback_leaves": st["fallback_leaves"], "avg_window": st["avg_window"],
                           "max_window": st["max_window"], "train_loss": st["train_loss"],
                           "index_bytes": st["memory_bytes"], "layout": st["layout"], "leaves": st["leaves"],
                           "avg_lines": st["avg_lines"], "max_lines": st["max_lines"],
                           "lds_levels": st["lds_levels"]}
            idx.close()
            log(f"C4 {dname} {label}: {m / ms * 1e-6:.2f} G lookups/s match={ok} fallback={st['fallback_keys']} "
                f"layout={st['layout']} leaves={st['leaves']} avg_lines={st['avg_lines']:.3f}")
        bt = rows["btree"]["lookups_per_s"]
        for label in ("learned_plr", "learned_mlp", "learned_plr_column", "learned_mlp_column"):
            rows[label]["vs_btree"] = rows[label]["lookups_per_s"] / bt
        c4["distributions"][dname] = rows
        del dk4, dp4, k4, e4
    res["C4_learned_misses"] = c4

    res["C3_partitioned"] = partitioned_leg(gj, G, stream, keys64, d_keys64, d_out)
    res["C5_join_aggregate"] = join_aggregate_leg(gj, stream, cpu=cpu)
    res["host_placement"] = host_placement_leg(gj, stream, keys64, d_keys64, d_out)
    return res


def partitioned_leg(gj, G, stream, keys, d_keys, d_out):
    """C3: the radix-partitioned driver over the headline column, 2^28 probes, uniform and
    Zipf(1.0): fan-out 2^8 .. 2^16 (above 2^12 in two passes, the second segmented),
    windows 2^20 / 2^24 / 2^26, for the B-tree and the learned index."""
    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    m = d_out.count
    out = {"workload": "2^26 u64 keys, 2^28 probes, radix-partitioned in windows, row ids in probe order",
           "rows": {}}
    for pdist, pname, seed in ((_lib.PROBES_UNIFORM, "uniform", 0x21), (_lib.PROBES_ZIPF, "zipf", 0x26)):
        probes, exp = gj.gen_probes(keys, m, pdist, 1.0, 0.0, 63, seed)
        dp = gj.DeviceArray.from_numpy(probes, stream)
        del probes
        for kind in ("btree", "learned_plr"):
            idx = gj.Index(kind, d_keys, host_column=keys, stream=stream)
            grid = [(8, 1, 24), (10, 1, 24), (12, 1, 24), (12, 2, 24), (14, 2, 24), (16, 2, 24), (10, 1, 20),
                    (10, 1, 26)] if pname == "uniform" else [(10, 1, 24), (16, 2, 24)]
            for bits, passes, wlog in grid:
                lp = gj.lookup_params("partitioning", radix_bits=bits, passes=passes, window=1 << wlog)
                ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=lp, stream=stream), stream, 2, 1)
                ok = bool(np.array_equal(d_out.numpy(m), exp))
                label = f"{pname}_{kind}_fan{1 << bits}_p{passes}_w2^{wlog}"
                out["rows"][label] = {"lookups_per_s": m / ms * 1e3,
                                      "gather_roofline_frac": m / ms * 1e3 / G if G else None, "rowids_match": ok}
                log(f"C3 {label}: {m / ms * 1e-6:.2f} G lookups/s match={ok}")
            assert idx.stats()["dropped_tuples"] == 0
            idx.close()
        del dp, exp
    return out


def host_placement_leg(gj, stream, keys, d_keys, d_out, mh=1 << 24):
    """The reference's placement: index structures (and the column) in page-locked host
    memory read over the host link, plain against partitioned lookups (fan-out 4096, one
    window).  2^26 u64 keys, 2^24 uniform probes in HBM."""
    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    probes, exp = gj.gen_probes(keys, mh, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x25)
    dp = gj.DeviceArray.from_numpy(probes, stream)
    hcol = gj.HostArray.from_numpy(keys)
    out = {"workload": "2^26 u64 keys, 2^24 uniform probes (in HBM); index structures in page-locked host memory",
           "rows": {}}
    # learned_plr in the default layout (the column layout when placed in host memory) and,
    # for contrast, in the lines layout, whose key copy then sits in host memory too
    for col_name, col in (("column_in_hbm", d_keys), ("column_in_host", hcol)):
        for kind, name, extra in (("btree", "btree", {}), ("learned_plr", "learned_plr", {}),
                                  ("learned_plr", "learned_plr_lines", {"learned_layout": _lib.LAYOUT_LINES})):
            idx = gj.Index(kind, col, host_column=keys, stream=stream,
                           params=gj.index_params(placement=_lib.PLACE_HOST, **extra))
            st = idx.stats()
            for appr, kw in (("plain", {}), ("partitioning", {"radix_bits": 12, "window": mh})):
                lp = gj.lookup_params(appr, **kw)
                ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=mh, params=lp, stream=stream), stream, 2, 1)
                ok = bool(np.array_equal(d_out.numpy(mh), exp))
                label = f"{name}_{appr}_{col_name}"
                out["rows"][label] = {"lookups_per_s": mh / ms * 1e3, "rowids_match": ok, "layout": st["layout"],
                                      "index_bytes_in_host": st["memory_bytes"]}
                log(f"host placement {label}: {mh / ms * 1e-6:.3f} G lookups/s match={ok}")
            idx.close()
    hcol.free()
    return out


def join_aggregate_leg(gj, stream, cpu=True, fact_rows=60_000_000, reps=6):
    """C5: 2e6-row dimension indexed by the learned model, 6e7-row fact table, 30-day
    window; one batch and log-uniform batches through launches / hipGraph / persistent
    kernel.  Repetitions alternate between two independently drawn copies of the fact
    table (2 x 1.02 GB), so no repetition finds its rows in the 256 MB MALL."""
    import oracle

    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    cores = os.cpu_count() or 1
    dim_rows = 2_000_000
    key, flag = gj.gen_dimension(dim_rows, 0x41)
    skey, sflag = gj.sort_dimension(key, flag)
    lo, hi = 1200, 1230
    copies = []
    for seed in (0x42, 0x44):
        cols = gj.gen_fact(fact_rows, dim_rows, 2557, 90000, 10500000, 10, seed)
        want = oracle.join_aggregate(skey, sflag, *cols, lo, hi, threads=cores)
        n_qual = int(np.count_nonzero((cols[1] >= lo) & (cols[1] < hi)))
        copies.append({"host": cols, "dev": [gj.DeviceArray.from_numpy(a, stream) for a in cols], "want": want,
                       "n_qual": n_qual})
    dflag = gj.DeviceArray.from_numpy(sflag, stream)
    dkey = gj.DeviceArray.from_numpy(skey, stream)
    rng = np.random.default_rng(0x43)
    sizes, total = [], 0
    while total < fact_rows:
        s = min(int(np.exp(rng.uniform(np.log(1e3), np.log(1e7)))), fact_rows - total)
        sizes.append(s)
        total += s
    out = {"workload": f"dimension {dim_rows} rows (dense shuffled int32 keys, promo p=1/3), fact {fact_rows} rows, "
                       f"30-day window (~1.2% selectivity), exact int64 sums; repetitions alternate two fact copies",
           "batches": len(sizes), "one_batch": {}, "plans": {}}
    res = gj.DeviceArray(4, np.int64)
    for kind in ("learned_plr", "btree", "hash"):
        idx = gj.Index(kind, dkey, host_column=skey, stream=stream)
        for strategy in (("lane_refill", "plain", "partitioned", "sorted", "lookup_first") if kind == "learned_plr"
                         else ("lane_refill",)):
            ok = True
            for c in copies:  # warm-up and check on both copies
                res.fill(0, stream)
                gj.join_aggregate(idx, *c["dev"], dflag, lo, hi, res, stream=stream, strategy=strategy)
                stream.synchronize()
                ok &= tuple(int(v) for v in res.numpy()[:3]) == c["want"]
            e = [gj.Event() for _ in range(reps + 1)]
            stream.synchronize()
            e[0].record(stream)
            for r in range(reps):
                res.fill(0, stream)
                gj.join_aggregate(idx, *copies[r % 2]["dev"], dflag, lo, hi, res, stream=stream, strategy=strategy)
                e[r + 1].record(stream)
            stream.synchronize()
            ms = e[0].elapsed_ms(e[reps]) / reps
            # bytes the kernels must read: every row's date, each qualifying row's fkey, price
            # and discount, each match's flag (late materialisation); and the survey's
            # definition, all four fact columns per row plus an 8 B gather and the flag per
            # qualifying row
            alg = fact_rows * 4 + copies[0]["n_qual"] * 13 + copies[0]["want"][2]
            alg_all = fact_rows * 17 + copies[0]["n_qual"] * 9
            label = kind if strategy == "lane_refill" else f"{kind}_{strategy}"
            out["one_batch"][label] = {"rows_per_s": fact_rows / ms * 1e3, "ms": ms, "sums_match": ok,
                                       "strategy": strategy, "algorithmic_GBps": alg / (ms * 1e-3) / 1e9,
                                       "all_columns_GBps": alg_all / (ms * 1e-3) / 1e9}
            log(f"C5 one batch {label}: {fact_rows / ms * 1e-6:.0f} G rows/s match={ok}")
        if kind != "learned_plr":
            idx.close()
            continue
        results = gj.DeviceArray(4 * len(sizes), np.int64)
        for mode, mname in ((_lib.BATCH_LAUNCH, "launch"), (_lib.BATCH_GRAPH, "hipgraph"),
                            (_lib.BATCH_PERSISTENT, "persistent")):
            plans = []
            for c in copies:
                batches, off = [], 0
                for s in sizes:
                    batches.append(tuple(col.offset(off) for col in c["dev"]) + (s,))
                    off += s
                plans.append(gj.AggregatePlan(idx, batches, dflag, lo, hi, results, mode, stream=stream))
            ok = True
            for p, c in zip(plans, copies):
                p.run(stream)
                stream.synchronize()
                sums = results.numpy().reshape(-1, 4)[:, :3].sum(axis=0)
                ok &= tuple(int(v) for v in sums) == c["want"]
            stream.synchronize()
            e[0].record(stream)
            for r in range(reps):
                plans[r % 2].run(stream)
                e[r + 1].record(stream)
            stream.synchronize()
            ms = e[0].elapsed_ms(e[reps]) / reps
            out["plans"][mname] = {"rows_per_s": fact_rows / ms * 1e3, "ms": ms, "sums_match": ok}
            log(f"C5 plan {mname}: {fact_rows / ms * 1e-6:.0f} G rows/s match={ok}")
            for p in plans:
                p.close()
        idx.close()
    if cpu:
        c = copies[0]["host"]
        sec, got = _cpu_median(lambda: oracle.join_aggregate(skey, sflag, *c, lo, hi, threads=cores))
        out["cpu_baseline"] = {"rows_per_s": fact_rows / sec, "cores": cores, "kind": "port",
                               "what": "the oracle's restated Q14 loop (filter, lower_bound, int64 sums), OpenMP, "
                                       "median of 10 after 3 warm-ups",
                               "sums_match": tuple(got) == copies[0]["want"]}
    return out
