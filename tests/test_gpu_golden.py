"""GPU results against the committed golden fixtures (tests/golden/, written by make_golden.py)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = ["binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash"]
FIXTURES = ["lookup_u64_uniform", "lookup_u32_uniform", "lookup_u64_lognormal", "lookup_u64_clustered"]


@pytest.mark.parametrize("fixture", FIXTURES)
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("approach", ["plain", "partitioning"])
def test_lookup_fixture(gj, gpu, fixture, kind, approach):
    g = np.load(os.path.join(GOLD, fixture + ".npz"), allow_pickle=False)
    keys, probes, want = g["keys"], g["probes"], g["rowids"]
    if kind == "harmonia" and keys.size % 32:
        pytest.skip("harmonia needs a multiple of 32 keys")
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_leaf_bits=8, mlp_steps=100))
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, approach=approach, radix_bits=6, window=5000)
    gj.synchronize()
    assert np.array_equal(out.numpy(), want)


@pytest.mark.parametrize("kind", ["btree", "learned_plr", "radix_spline"])
def test_join_aggregate_fixture(gj, gpu, kind):
    g = np.load(os.path.join(GOLD, "join_agg.npz"), allow_pickle=False)
    lo, hi = (int(v) for v in g["window"])
    col = gj.DeviceArray.from_numpy(g["dim_key"])
    idx = gj.Index(kind, col, host_column=g["dim_key"])
    d = [gj.DeviceArray.from_numpy(g[c]) for c in ("fkey", "date", "price", "disc")]
    res = gj.DeviceArray(4, np.int64)
    res.fill(0)
    gj.join_aggregate(idx, *d, gj.DeviceArray.from_numpy(g["dim_flag"]), lo, hi, res)
    gj.synchronize()
    assert np.array_equal(res.numpy()[:3], g["sums"])
