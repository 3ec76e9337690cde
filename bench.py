#!/usr/bin/env python3
"""Index-join probe benchmark on one MI355X.

Workload (the configuration the metric is quoted on): a sorted build column of
2^26 unique uint64 keys drawn uniformly from [0, 2^63), row id = position;
2^28 uint64 probe keys drawn uniformly with replacement from the build keys
(100 % hit).  One step = one batched lookup of all 2^28 probes through the C
ABI (gj_index_lookup), inputs resident in HBM before timing starts.

Reported:
  value            lookups/s of the headline index (learned when available, else the B-tree)
  roofline         HBM roofline of the headline lookup kernel: 20 algorithmic bytes per lookup
                   (8 B probe key + 8 B indexed key + 4 B row id) / average kernel time, vs 8 TB/s
  gather_roofline  G = independent random 8 B gathers per second measured on the same GPU in the same
                   process (gj_gather), and lookups/s / G per index
  hbm_traffic      measured HBM bytes per lookup launch (rocprofv3 counters, committed under profiles/)
                   over the live average kernel time: how close the kernel runs to the 8 TB/s peak
  l2               L2 hit rate and L2 misses per lookup of the headline kernel (rocprofv3 counters,
                   committed under profiles/): lookups/s = the gather probe's miss rate / misses per lookup
  cpu_baseline     the oracle's OpenMP equality-checked lower_bound (port of binary_search_index) on a
                   bounded probe sample, on this host's cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
BYTES_PER_LOOKUP = 20  # 8 B probe + 8 B indexed key + 4 B row id


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup, stream, gj):
    """Runs fn() warmup + steps times; returns (per-step ms list, total ms) from events on `stream`."""
    for _ in range(warmup):
        fn()
    stream.synchronize()
    ev = [gj.Event() for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        fn()
        ev[i + 1].record(stream)
    stream.synchronize()
    per = [ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)]
    return per, ev[0].elapsed_ms(ev[steps])


def _profile_values(file_name: str, kernel_tag: str):
    """Numbers after `<tag>` on its line of a committed profiles/ table, or None."""
    path = os.path.join(ROOT, "profiles", file_name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == kernel_tag:
                try:
                    return [float(v) for v in parts[1:]]
                except ValueError:
                    return None
    return None


def load_traffic(kernel_tag: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    committed as profiles/traffic.txt lines `<tag> <bytes_per_launch>`."""
    v = _profile_values("traffic.txt", kernel_tag)
    return v[0] if v else None


def load_l2(kernel_tag: str):
    """L2 hit rate and misses per lookup measured with rocprofv3 --pmc (TCC_HIT_sum, TCC_MISS_sum),
    committed as profiles/l2.txt lines `<tag> <hit_rate> <misses_per_lookup>`."""
    v = _profile_values("l2.txt", kernel_tag)
    return {"hit_rate": v[0], "misses_per_lookup": v[1],
            "source": "rocprofv3 TCC_HIT_sum / TCC_MISS_sum, profiles/l2.txt"} if v and len(v) >= 2 else None


def _index_params_str(p) -> str:
    return (f"lanes={p.harmonia_group_lanes} btree_node={p.btree_node_bytes} btree_lds={p.btree_lds_bytes} "
            f"rs_bits={p.rs_radix_bits} rs_lds_bits={p.rs_lds_bits} "
            f"rs_err={p.rs_max_error} layout={p.learned_layout} line_keys={p.learned_line_keys} "
            f"leaf_bits={p.learned_leaf_bits} max_err={p.learned_max_error}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-n", type=int, default=26)
    ap.add_argument("--log2-m", type=int, default=28)
    ap.add_argument("--indexes", default="binary_search,btree,harmonia,radix_spline,learned_plr,learned_mlp,hash")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-log2", type=int, default=23)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other configurations' legs (u32, Zipf, learned with misses, join-aggregate)")
    ap.add_argument("--profile-index", default="", help="run only this index (for rocprofv3 runs)")
    ap.add_argument("--approach", default="auto", help="auto (plain + autotuned), plain, bws, partitioning or hj")
    ap.add_argument("--tune-log2", type=int, default=24, help="probe sample size of the autotuner")
    ap.add_argument("--tune-index", action="store_true",
                    help="pick each index's own parameters with the index autotuner before building it")
    ap.add_argument("--only", choices=("C3", "sweep", "C5", "host", "host_sweep", "other"), default="",
                    help="run one leg of the other configurations alone and print its JSON (for rocprofv3 runs): "
                         "C3 the partitioned driver, sweep the index-size sweep, C5 the join-aggregate, host the "
                         "host-memory placement, host_sweep columns of 2^26..2^31 keys in host memory (not in the "
                         "default run), other sorted/dense probes, the hash-join driver and lookup plans")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "the probe path runs on one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)

    import gpujoin_amd as gj
    from gpujoin_amd import _lib

    info = gj.device_info()
    log("device", info)
    if args.only == "C5":
        import bench_suite as suite
        print(json.dumps({"C5_join_aggregate": suite.join_aggregate_leg(gj, gj.Stream(), cpu=not args.no_cpu)}))
        return
    if args.only == "host_sweep":
        import bench_suite as suite
        print(json.dumps({"host_sweep": suite.host_sweep_leg(gj, gj.Stream())}))
        return
    n = 1 << args.log2_n
    m = 1 << args.log2_m
    threads = 0
    t0 = time.time()
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x11, threads)
    probes, expected = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x12, threads)
    log(f"generated n={n} m={m} in {time.time() - t0:.1f}s")
    stream = gj.Stream()
    d_keys = gj.DeviceArray.from_numpy(keys, stream)
    d_probes = gj.DeviceArray.from_numpy(probes, stream)
    d_out = gj.DeviceArray(m, np.uint32)
    stream.synchronize()
    if args.only:
        import bench_suite as suite
        legs = {"C3": lambda: suite.partitioned_leg(gj, None, stream, keys, d_keys, d_out),
                "sweep": lambda: suite.index_sweep_leg(gj, None, stream, d_out),
                "host": lambda: suite.host_placement_leg(gj, stream, keys, d_keys, d_out),
                "other": lambda: suite.other_leg(gj, None, stream, keys, d_keys, d_out)}
        leg = legs[args.only]()
        print(json.dumps({args.only: leg}))
        return

    results = {}

    # ---- gather roofline G --------------------------------------------------
    best_g = None
    for items in (1, 2, 4, 8):
        fn = lambda: gj.gather(d_keys, d_probes, d_out, m, 256, items, stream)
        per, _ = time_steps(fn, max(3, args.steps // 2), 1, stream, gj)
        ms = statistics.median(per)
        rate = m / (ms * 1e-3)
        log(f"gather items={items}: {ms:.3f} ms  {rate / 1e9:.2f} G gathers/s")
        if best_g is None or rate > best_g["rate"]:
            best_g = {"rate": rate, "ms": ms, "items": items}
    G = best_g["rate"]

    # streaming copy bandwidth (1 GiB)
    cb = 1 << 30
    src = gj.DeviceArray(cb // 8, np.uint64)
    dst = gj.DeviceArray(cb // 8, np.uint64)
    src.fill(1, stream)
    per, _ = time_steps(lambda: gj.stream_copy(src, dst, cb, stream), 5, 2, stream, gj)
    copy_gbps = 2 * cb / (statistics.median(per) * 1e-3) / 1e9
    log(f"stream copy: {copy_gbps:.0f} GB/s (read + write)")
    # streaming read of the same 1 GiB (the bound of read-only kernels: the join-aggregate)
    xo = gj.DeviceArray(1, np.uint64)
    per, _ = time_steps(lambda: gj.stream_read(src, cb, xo, stream), 5, 2, stream, gj)
    read_gbps = cb / (statistics.median(per) * 1e-3) / 1e9
    del src, dst, xo
    log(f"stream read: {read_gbps:.0f} GB/s")

    # dependent pointer chase over 512 MiB (latency)
    nxt_n = 1 << 27
    perm = np.random.default_rng(7).permutation(nxt_n).astype(np.uint32)
    nxt = np.empty(nxt_n, dtype=np.uint32)
    nxt[perm] = np.roll(perm, -1)  # one cycle through all slots
    d_next = gj.DeviceArray.from_numpy(nxt, stream)
    chase_out = gj.DeviceArray(1, np.uint32)
    steps_chase = 100000
    gj.pointer_chase(d_next, 1, 1, chase_out, stream)
    # every launch continues the chain where the previous one stopped, so no launch
    # re-walks lines that L2 or the MALL may still hold (the cycle spans 2^27 slots)
    per, _ = time_steps(lambda: gj.pointer_chase(d_next, 1, steps_chase, chase_out, stream, start=chase_out), 3, 1,
                        stream, gj)
    chase_ns = statistics.median(per) * 1e6 / steps_chase
    del d_next, nxt, perm
    log(f"pointer chase: {chase_ns:.0f} ns per dependent access")

    # random 8 B gathers from page-locked host memory (the reference's random_access
    # setting: a zero-copy host array read over the interconnect)
    host_arr = gj.HostArray.from_numpy(keys)
    mh = min(m, 1 << 24)
    per, _ = time_steps(lambda: gj.gather(host_arr, d_probes, d_out, mh, 256, 8, stream), 3, 1, stream, gj)
    G_host = mh / (statistics.median(per) * 1e-3)
    host_arr.free()
    del host_arr
    log(f"gather from host memory: {G_host / 1e9:.3f} G gathers/s")

    # ---- index lookups ------------------------------------------------------
    names = [args.profile_index] if args.profile_index else [s for s in args.indexes.split(",") if s]
    approach_names = {0: "plain", 1: "bws", 2: "partitioning", 3: "hj"}
    tune_m = min(m, 1 << args.tune_log2)
    for name in names:
        t_build = time.time()
        try:
            iparams = None
            if args.tune_index:
                iparams, irate = gj.autotune_index(name, d_keys, d_probes, d_out, host_column=keys, m=tune_m,
                                                   stream=stream)
                log(f"{name}: index autotune on 2^{args.tune_log2} probes -> {_index_params_str(iparams)} "
                    f"({irate / 1e9:.2f} G/s)")
            idx = gj.Index(name, d_keys, host_column=keys, params=iparams, stream=stream)
        except gj.GjError as e:
            log(f"{name}: not built ({e})")
            continue
        build_s = time.time() - t_build
        st = idx.stats()
        configs = [("plain", gj.lookup_params("plain"))]
        if args.approach == "auto":
            best, rate_s = gj.autotune(idx, d_probes, d_out, m=tune_m, stream=stream)
            log(f"{name}: autotune on 2^{args.tune_log2} probes -> {approach_names[best.approach]} "
                f"bits={best.radix_bits} window={best.window} passes={best.passes} ({rate_s / 1e9:.2f} G/s)")
            if best.approach != 0 or best.block_size != 256:
                configs.append(("tuned", best))
        elif args.approach != "plain":
            configs = [(args.approach, gj.lookup_params(args.approach))]
        entry = {"index_bytes": st["memory_bytes"], "depth": st["depth"], "lds_levels": st["lds_levels"],
                 "build_s": build_s}
        if iparams is not None:
            entry["index_params"] = _index_params_str(iparams)
        for k in ("spline_points", "leaves", "fallback_leaves", "fallback_keys", "max_window", "avg_window",
                  "train_loss", "layout", "lines", "avg_lines", "max_lines"):
            if st[k]:
                entry[k] = st[k]
        best_rate = 0.0
        for label, lp in configs:
            fn = lambda: idx.lookup(d_probes, d_out, m=m, params=lp, stream=stream)
            per, total = time_steps(fn, args.steps, args.warmup, stream, gj)
            got = d_out.numpy(m)
            ok = bool(np.array_equal(got, expected))
            avg = sum(per) / len(per)
            rate = m * args.steps / (total * 1e-3)
            res = {"approach": approach_names[lp.approach], "radix_bits": lp.radix_bits, "window": lp.window,
                   "passes": lp.passes, "block": lp.block_size, "lookups_per_s": rate,
                   "ms_per_step": total / args.steps, "avg_launch_ms": avg, "gather_roofline_frac": rate / G,
                   "rowids_match": ok}
            entry[label] = res
            log(f"{name} [{label} {res['approach']}]: {rate / 1e9:.3f} G lookups/s, {avg:.3f} ms/step, "
                f"{rate / G * 100:.1f}% of G, match={ok}")
            if not ok:
                log(f"{name}: ROW ID MISMATCH ({int(np.count_nonzero(got != expected))} probes)")
            if ok and rate > best_rate:
                best_rate = rate
                entry.update({"lookups_per_s": rate, "ms_per_step": total / args.steps, "avg_launch_ms": avg,
                              "best": label, "gather_roofline_frac": rate / G, "rowids_match": True})
        entry.setdefault("rowids_match", False)
        results[name] = entry
        idx.close()

    # ---- CPU baseline (oracle port, bounded sample) ---------------------------
    # 3 warm-ups + 10 repetitions, median (the reference's measure(), src/measuring.hpp:12),
    # for the equality-checked lower_bound and the reference's B-tree descent
    cpu = None
    if not args.no_cpu:
        import oracle
        ms_n = 1 << args.cpu_sample_log2
        sample = probes[:ms_n]
        cores = os.cpu_count() or 1

        def cpu_leg(fn):
            for _ in range(3):
                out_cpu = fn()
            times = []
            for _ in range(10):
                t = time.perf_counter()
                out_cpu = fn()
                times.append(time.perf_counter() - t)
            return ms_n / statistics.median(times), bool(np.array_equal(out_cpu, expected[:ms_n]))

        lb_rate, lb_ok = cpu_leg(lambda: oracle.lookup_lower_bound(keys, sample, threads=cores))
        tree = oracle.BTree(keys, 4)
        bt_rate, bt_ok = cpu_leg(lambda: tree.lookup(sample, threads=cores))
        del tree
        cpu = {"value": lb_rate, "unit": "lookups/s", "cores": cores, "kind": "port",
               "sample": f"2^{args.cpu_sample_log2} of the 2^{args.log2_m} probes against the full 2^{args.log2_n}-key "
                         f"column, equality-checked lower_bound, OpenMP static schedule, median of 10 after 3 "
                         f"warm-ups; row ids match: {lb_ok}",
               "btree_descent": {"value": bt_rate, "unit": "lookups/s", "rowids_match": bt_ok,
                                 "what": "the reference's clustered B+-tree (4 KiB pages) restated in the oracle, "
                                         "same sample, threads and repetitions"}}
        log(f"cpu baseline: lower_bound {lb_rate / 1e6:.1f} M lookups/s, B-tree {bt_rate / 1e6:.1f} M lookups/s "
            f"on {cores} threads")

    # the headline is the faster learned index (both roots are learned models), else the next in order
    learned = [h for h in ("learned_plr", "learned_mlp") if h in results and results[h]["rowids_match"]]
    head_order = ["radix_spline", "btree", "harmonia", "binary_search"]
    head = max(learned, key=lambda h: results[h]["lookups_per_s"]) if learned else \
        next((h for h in head_order if h in results and results[h]["rowids_match"]), None)
    if head is None and results:
        head = next(iter(results))
    r = results.get(head, {})
    avg_s = r.get("avg_launch_ms", float("nan")) * 1e-3
    achieved = m * BYTES_PER_LOOKUP / avg_s / 1e9 if r else None
    traffic = load_traffic(f"{head}_u64_2^{args.log2_n}_2^{args.log2_m}") if head else None
    line = {
        "metric": "probe lookups/s + % gather roofline, learned vs B-tree, 2^26-key index, 1 GPU",
        "value": r.get("lookups_per_s"),
        "unit": "lookups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r.get("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded): uniform unique u64 build keys, uniform probes",
        "config": {"workload": f"index probe: 2^{args.log2_n} sorted unique u64 keys in [0, 2^63), "
                               f"2^{args.log2_m} uniform u64 probes (100% hit), u32 row ids",
                   "index": head, "approach": "plain", "device": info["name"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": (achieved / PEAK_HBM_GBPS) if achieved else None, "traffic": traffic},
        "gather_roofline": {"G_gathers_per_s": G, "items_per_thread": best_g["items"],
                            "frac": (r.get("lookups_per_s", 0) / G) if r else None},
        "cpu_baseline": cpu,
        "indexes": results,
        "learned_vs_btree": (results[head]["lookups_per_s"] / results["btree"]["lookups_per_s"])
        if head in ("learned_mlp", "learned_plr", "radix_spline") and "btree" in results else None,
        "hbm_traffic": {"bytes_per_launch": traffic, "bytes_per_lookup": traffic / m if traffic else None,
                        "GBps": traffic / avg_s / 1e9 if (traffic and r) else None,
                        "frac_of_peak": traffic / avg_s / 1e9 / PEAK_HBM_GBPS if (traffic and r) else None,
                        "source": "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE (copy-calibrated), profiles/traffic.txt"},
        "l2": load_l2(f"{head}_u64_2^{args.log2_n}_2^{args.log2_m}") if head else None,
        "stream_copy_GBps": copy_gbps,
        "stream_read_GBps": read_gbps,
        "pointer_chase_ns": chase_ns,
        "host_gather_per_s": G_host,
    }
    if not args.no_extras:
        import bench_suite as suite
        del d_probes, probes, expected
        line["configs"] = suite.extras(gj, G, stream, keys, d_keys, d_out, cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
