#!/usr/bin/env python3
"""Writes the committed golden fixtures of tests/golden/ (run from the repository root).

The reference ships no data fixtures and its only checkable tests are the
Harmonia host-lookup KATs (test/test_harmonia.cu:18-55, 115-123;
test/main.cu:144-164) and the Zipf sampler (src/zipf.hpp), which is compiled
from the reference tree into oracle/_ref.  These fixtures hold:
  * lookup_*.npz     seeded columns/probes and the oracle's row ids (lower_bound + equality)
  * harmonia_kat.npz the Harmonia arrays (inner keys, children) for the reference KAT shapes
  * rs_u64.npz       RadixSpline knots and radix table of a seeded column (the builder's output)
  * join_agg.npz     a small fact/dimension pair with its exact int64 sums
  * zipf_ref.npz     draws of the reference's own zipf_distribution (std::mt19937, fixed seeds)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import gpujoin_amd as gj  # noqa: E402  (host-side generators only; no GPU needed)
import oracle  # noqa: E402
from gpujoin_amd import _lib  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def lookups():
    for kb, dist, name in [(8, _lib.KEYS_UNIFORM, "lookup_u64_uniform"), (4, _lib.KEYS_UNIFORM, "lookup_u32_uniform"),
                           (8, _lib.KEYS_LOGNORMAL, "lookup_u64_lognormal"),
                           (8, _lib.KEYS_CLUSTERED, "lookup_u64_clustered")]:
        n = 4096 if dist != _lib.KEYS_UNIFORM else 8192
        keys = gj.gen_build_keys(n, kb, dist, 0, 0x1001 + kb + dist)
        probes, _ = gj.gen_probes(keys, 16384, _lib.PROBES_UNIFORM, 1.0, 0.1, 0, 0x2001 + kb + dist)
        zipf, _ = gj.gen_probes(keys, 4096, _lib.PROBES_ZIPF, 1.0, 0.0, 0, 0x3001 + kb + dist)
        probes = np.concatenate([probes, zipf])
        rowids = oracle.lookup_lower_bound(keys, probes)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), keys=keys, probes=probes, rowids=rowids)


def harmonia():
    out = {}
    for n, f in [(280, 9), (2000, 9), (144, 5), (32 * 34, 33), (32 * 1100, 33)]:
        keys = np.arange(n, dtype=np.uint32) * 3 + 1
        h = oracle.Harmonia(keys, f)
        k, c = h.arrays()
        out[f"keys_{n}_{f}"] = k
        out[f"children_{n}_{f}"] = c
        out[f"depth_{n}_{f}"] = np.array([h.depth()])
    np.savez_compressed(os.path.join(OUT, "harmonia_kat.npz"), **out)


def radix_spline():
    keys = gj.gen_build_keys(20000, 8, _lib.KEYS_LOGNORMAL, 0, 0x4001)
    rs = oracle.RadixSpline(keys, 10, 8)
    px, py, radix = rs.arrays()
    np.savez_compressed(os.path.join(OUT, "rs_u64.npz"), keys=keys, knot_x=px, knot_y=py, radix=radix,
                        shift=np.array([rs.shift_bits()]))


def join_agg():
    key, flag = gj.gen_dimension(3000, 0x5001)
    skey, sflag = gj.sort_dimension(key, flag)
    fkey, date, price, disc = gj.gen_fact(50000, 4000, 2557, 90000, 10500000, 10, 0x5002)
    num, den, mt = oracle.join_aggregate(skey, sflag, fkey, date, price, disc, 100, 400)
    np.savez_compressed(os.path.join(OUT, "join_agg.npz"), dim_key=skey, dim_flag=sflag, fkey=fkey, date=date,
                        price=price, disc=disc, window=np.array([100, 400]), sums=np.array([num, den, mt]))


def zipf():
    if not oracle.ref_zipf_available():
        print("oracle/_ref/libref_zipf.so absent: zipf_ref.npz not regenerated")
        return
    out = {}
    for n, q, seed in [(100, 1.0, 1), (2 ** 26 - 1, 1.0, 7), (1000, 0.5, 3), (1000, 1.25, 5), (50, 2.0, 11)]:
        out[f"draws_{n}_{q}_{seed}"] = oracle.ref_zipf_mt19937(n, q, seed, 2000)
    np.savez_compressed(os.path.join(OUT, "zipf_ref.npz"), **out)


if __name__ == "__main__":
    lookups()
    harmonia()
    radix_spline()
    join_agg()
    zipf()
    print("golden fixtures written to", OUT)
