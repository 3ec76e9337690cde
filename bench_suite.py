"""The legs of the default bench run beyond the headline (bench.py puts their results
under "configs" in its JSON line; `bench.py --only C3|sweep|C5|host|host_sweep|other` runs
one leg alone).

  C1_u32             2^20 unique u32 keys, 2^24 probes: row ids + u64 checksum, every index
  C2b_zipf           the headline's 2^26 u64 keys, 2^28 Zipf(1.0) probes, every index
  C3_partitioned     fan-out 2^8..2^16 (one and two passes), windows 2^16..2^26, uniform and Zipf
  C3_index_sweep     2^20..2^30 unique u64 keys (up to 8 GiB), plain and partitioned
  C4_learned_misses  uniform / lognormal / clustered keys, 10 % misses, RadixSpline err 8/32/64,
                     learned in both layouts, fallback keys counted
  C5_join_aggregate  2e6-row dimension, 6e7 fact rows, every strategy, the three batch modes
  host_placement     index structures (and the column) in page-locked host memory
  host_sweep         (not in the default run) dense columns of 2^26..2^31 keys in host memory,
                     partitioned lookups, after the reference's TLB study
  other              sorted probes, dense keys, the hash-join driver, the B-tree's node sizes,
                     lookup plans

Every measurement checks its row ids or sums against the generator (or the oracle for
the aggregates).
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


# ---- the other configurations' legs of the default bench run ------------------------------

_APPROACHES = {0: "plain", 1: "bws", 2: "partitioning", 3: "hj"}


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
    """Every leg in the module docstring, measured in the default run, every result checked."""
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
    res["C3_index_sweep"] = index_sweep_leg(gj, G, stream, d_out)
    res["other"] = other_leg(gj, G, stream, keys64, d_keys64, d_out)
    res["C5_join_aggregate"] = join_aggregate_leg(gj, stream, cpu=cpu)
    res["host_placement"] = host_placement_leg(gj, stream, keys64, d_keys64, d_out)
    return res


def partitioned_leg(gj, G, stream, keys, d_keys, d_out):
    """C3: the radix-partitioned driver over the headline column, 2^28 probes, uniform and
    Zipf(1.0): fan-out 2^8 .. 2^16 (above 2^12 in two passes, the second segmented),
    windows 2^16 .. 2^26, for the B-tree and the learned index."""
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
            # fan-out 2^8 / 2^10 / 2^12 (one and two passes) over windows of 2^16 .. 2^22 probes;
            # 2^14 and 2^16 (two passes, the second segmented); larger windows
            grid = ([(b, p, w) for b, p in ((8, 1), (10, 1), (12, 1), (12, 2)) for w in (16, 18, 20, 22)] +
                    [(14, 2, 22), (16, 2, 22), (10, 1, 24), (10, 1, 26)]) if pname == "uniform" else \
                [(10, 1, 22), (12, 2, 22), (16, 2, 22)]
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


def index_sweep_leg(gj, G, stream, d_out):
    """C3's index-size sweep: 2^20 .. 2^30 unique u64 keys (the largest column 8 GiB), 2^28
    uniform probes, the B-tree and the learned index through the plain driver and the
    partitioned one (fan-out 2^10, windows of 2^24)."""
    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    m = d_out.count
    out = {"workload": "2^20..2^30 unique u64 keys from [0, 2^63), 2^28 uniform probes (100% hit)", "rows": {}}
    for s_log in (20, 22, 24, 26, 28, 30):
        keys = gj.gen_build_keys(1 << s_log, 8, _lib.KEYS_UNIFORM, 63, 0x22 + s_log)
        probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x23 + s_log)
        dk = gj.DeviceArray.from_numpy(keys, stream)
        dp = gj.DeviceArray.from_numpy(probes, stream)
        del probes
        for kind in ("btree", "learned_plr"):
            idx = gj.Index(kind, dk, host_column=keys, stream=stream)
            st = idx.stats()
            for appr in ("plain", "partitioning"):
                lp = gj.lookup_params(appr)
                ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=lp, stream=stream), stream, 2, 1)
                ok = bool(np.array_equal(d_out.numpy(m), exp))
                label = f"2^{s_log}_{kind}_{appr}"
                out["rows"][label] = {"lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G if G
                                      else None, "rowids_match": ok, "index_bytes": st["memory_bytes"],
                                      "layout": st["layout"]}
                log(f"C3 sweep {label}: {m / ms * 1e-6:.2f} G lookups/s match={ok}")
            idx.close()
        del dk, dp, keys, exp
    return out


def other_leg(gj, G, stream, keys, d_keys, d_out):
    """Sorted probes (the reference's -s true), dense keys (its --dataset dense), the
    hash-join driver (a table built over the column on every call), the B-tree's node
    sizes, and uneven probe batches against one resident index through per-batch
    launches, a hipGraph and the persistent kernel.  The headline's 2^26 keys and 2^28
    probes."""
    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    m = d_out.count
    res = {"sorted_probes": {}, "dense_keys": {}, "hash_join_driver": {}, "lookup_plans": {}, "btree_node_bytes": {}}

    def run(section, label, idx, dp, exp, lp):
        ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=lp, stream=stream), stream, 3, 1)
        ok = bool(np.array_equal(d_out.numpy(m), exp))
        res[section][label] = {"lookups_per_s": m / ms * 1e3, "gather_roofline_frac": m / ms * 1e3 / G if G else None,
                               "rowids_match": ok}
        log(f"{section} {label}: {m / ms * 1e-6:.2f} G lookups/s match={ok}")

    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x17)
    order = np.argsort(probes, kind="stable")
    dp = gj.DeviceArray.from_numpy(probes[order], stream)
    exp_sorted = exp[order]
    del probes, exp, order
    for kind in ("btree", "harmonia", "radix_spline", "learned_plr"):
        idx = gj.Index(kind, d_keys, host_column=keys, stream=stream)
        run("sorted_probes", kind, idx, dp, exp_sorted, gj.lookup_params())
        idx.close()
    del exp_sorted

    dense = gj.gen_build_keys(keys.size, 8, _lib.KEYS_DENSE, 0, 0)
    dd = gj.DeviceArray.from_numpy(dense, stream)
    probes, exp = gj.gen_probes(dense, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 0, 0x18)
    dp.upload(probes)
    for kind in ("binary_search", "btree", "harmonia", "radix_spline", "learned_plr"):
        idx = gj.Index(kind, dd, host_column=dense, stream=stream)
        run("dense_keys", kind, idx, dp, exp, gj.lookup_params())
        idx.close()
    del dd, dense

    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x14)
    dp.upload(probes)
    idx = gj.Index("btree", d_keys, host_column=keys, stream=stream)
    run("hash_join_driver", "table_build_and_probe", idx, dp, exp, gj.lookup_params("hj"))
    idx.close()
    # the B-tree's node size (and leaf chunk): half a line, one line (default), two lines
    for nb in (64, 128, 256):
        idx = gj.Index("btree", d_keys, host_column=keys, stream=stream, params=gj.index_params(btree_node_bytes=nb))
        run("btree_node_bytes", str(nb), idx, dp, exp, gj.lookup_params())
        res["btree_node_bytes"][str(nb)]["depth"] = idx.stats()["depth"]
        idx.close()

    rng = np.random.default_rng(0x15)
    sizes, total = [], 0
    while total < m:
        sz = min(int(np.exp(rng.uniform(np.log(1e3), np.log(1e7)))), m - total)
        sizes.append(sz)
        total += sz
    idx = gj.Index("learned_plr", d_keys, host_column=keys, stream=stream)
    batches, off = [], 0
    for sz in sizes:
        batches.append((dp.offset(off), d_out.offset(off), sz))
        off += sz
    res["lookup_plans"]["batches"] = len(sizes)
    for mode, mname in ((_lib.BATCH_LAUNCH, "launch"), (_lib.BATCH_GRAPH, "hipgraph"),
                        (_lib.BATCH_PERSISTENT, "persistent")):
        plan = gj.LookupPlan(idx, batches, mode, stream=stream)
        ms = _timed(gj, lambda: plan.run(stream), stream, 3, 1)
        ok = bool(np.array_equal(d_out.numpy(m), exp))
        res["lookup_plans"][mname] = {"lookups_per_s": m / ms * 1e3, "rowids_match": ok}
        log(f"lookup plan {mname}: {m / ms * 1e-6:.2f} G lookups/s match={ok}")
        plan.close()
    ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=gj.lookup_params(), stream=stream), stream, 3, 1)
    res["lookup_plans"]["one_batch"] = {"lookups_per_s": m / ms * 1e3,
                                        "rowids_match": bool(np.array_equal(d_out.numpy(m), exp))}
    idx.close()
    del dp, probes, exp
    return res


def host_placement_leg(gj, stream, keys, d_keys, d_out, mh=1 << 24):
    """The reference's placement: index structures (and the column) in page-locked host
    memory read over the host link, plain lookups against the configuration the lookup
    autotuner picks on the same probes (in practice the partitioned driver; a fixed fan-out
    of 4096 in one pass measured 20 % slower, its scatter falling back to register tiles).
    2^26 u64 keys, 2^24 uniform probes in HBM."""
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
            tuned, _ = gj.autotune(idx, dp, d_out, m=mh, stream=stream)
            for appr, lp in (("plain", gj.lookup_params("plain")), ("tuned", tuned)):
                ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=mh, params=lp, stream=stream), stream, 2, 1)
                ok = bool(np.array_equal(d_out.numpy(mh), exp))
                label = f"{name}_{appr}_{col_name}"
                out["rows"][label] = {"lookups_per_s": mh / ms * 1e3, "rowids_match": ok, "layout": st["layout"],
                                      "index_bytes_in_host": st["memory_bytes"],
                                      "approach": _APPROACHES[lp.approach], "radix_bits": lp.radix_bits,
                                      "passes": lp.passes, "window": lp.window}
                log(f"host placement {label}: {mh / ms * 1e-6:.3f} G lookups/s match={ok} "
                    f"({_APPROACHES[lp.approach]}, 2^{lp.radix_bits} x {lp.passes} pass(es), window {lp.window})")
            idx.close()
    hcol.free()
    return out


def host_sweep_leg(gj, stream, sizes=(26, 28, 30, 31), m=1 << 26):
    """The host placement at scale, after the reference's TLB study
    (scripts/bench_tlb_misses.sh): dense keys (key = row id, its --dataset dense) in a
    column held in page-locked host memory, 2^26 keys (0.5 GiB) to 2^31 (16 GiB; row ids
    are u32, so a column stays below 2^32 keys), the index structures in host memory too,
    2^26 uniform probes in HBM through the partitioned driver (fan-out 2^12, one window) as
    the script runs them, and the learned index also through the plain driver.  Not part
    of the default run (`bench.py --only host_sweep`)."""
    log = lambda *a: print("[extras]", *a, file=sys.stderr, flush=True)
    rng = np.random.default_rng(0x27)
    out = {"workload": "dense u64 keys in page-locked host memory (2^26..2^31 keys), index structures in host "
                       "memory, 2^26 uniform probes in HBM", "rows": {}}
    d_out = gj.DeviceArray(m, np.uint32)
    chunk = 1 << 26
    for lg in sizes:
        n = 1 << lg
        hcol = gj.HostArray(n, np.uint64)
        view = hcol.view()
        for c in range(0, n, chunk):
            view[c:c + chunk] = np.arange(c, min(n, c + chunk), dtype=np.uint64)
        probes = rng.integers(0, n, size=m, dtype=np.uint64)
        exp = probes.astype(np.uint32)  # dense keys: key k sits at row k
        dp = gj.DeviceArray.from_numpy(probes, stream)
        for kind in ("btree", "learned_plr", "harmonia", "radix_spline", "binary_search"):
            t0 = time.time()
            idx = gj.Index(kind, hcol, host_column=view, stream=stream,
                           params=gj.index_params(placement=_lib.PLACE_HOST))
            build_s = time.time() - t0
            st = idx.stats()
            runs = [("partitioning", {"radix_bits": 12, "window": m})]
            if kind == "learned_plr":
                runs.append(("plain", {}))
            for appr, kw in runs:
                lp = gj.lookup_params(appr, **kw)
                ms = _timed(gj, lambda: idx.lookup(dp, d_out, m=m, params=lp, stream=stream), stream, 2, 1)
                ok = bool(np.array_equal(d_out.numpy(m), exp))
                label = f"2^{lg}_{kind}_{appr}"
                out["rows"][label] = {"lookups_per_s": m / ms * 1e3, "rowids_match": ok, "build_s": build_s,
                                      "column_bytes": 8 * n, "index_bytes_in_host": st["memory_bytes"]}
                log(f"host sweep {label}: {m / ms * 1e-6:.3f} G lookups/s match={ok} build {build_s:.1f}s")
            idx.close()
        del dp, probes, exp, view
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
            # and discount, each match's flag (late materialisation); and, for comparison, all
            # four fact columns per row plus an 8 B gather and the flag per qualifying row
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
