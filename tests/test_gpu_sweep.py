"""GPU parity across the index-size sweep: 2^24, 2^28 and 2^30 unique u64 keys
(the largest column is 8 GiB, past 2^32-byte offsets, where the reference's int /
unsigned indexing breaks: src/index_lookup_common.cuh:218-236,
src/index_lookup_partitioning.cu:445-447; sweep of scripts/bench_varying_index_size.sh).

For each size a 2^22-probe sample with 10 % misses runs through the plain driver
and the partitioned driver (fan-out 2^10 in one pass, 2^12 in two passes) for the
B-tree, the learned index, binary search and the hash table; hits must return the generator's rank
and misses must be absent from the host column (checked here by np.searchsorted).
One Harmonia index is built and probed at 2^30 keys.
"""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu

SIZES = [24, 28, 30]
M = 1 << 22


@pytest.fixture(scope="module", params=SIZES, ids=[f"2^{s}" for s in SIZES])
def column(request, gj, gpu):
    s = request.param
    n = 1 << s
    print(f"[case] sweep column 2^{s} keys ({(8 << s) / 2 ** 30:.2f} GiB)", flush=True)
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x22 + s)
    probes, exp = gj.gen_probes(keys, M, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x23 + s)
    # the generator's expectation, checked independently: a hit's rank holds the probe,
    # a miss is absent from the column
    hit = exp != gj.NOT_FOUND
    assert np.array_equal(keys[exp[hit]], probes[hit])
    pos = np.searchsorted(keys, probes[~hit])
    pos = np.minimum(pos, n - 1)
    assert not np.any(keys[pos] == probes[~hit])
    assert 0.08 < (~hit).mean() < 0.12
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(M, np.uint32)
    yield s, keys, col, dp, out, exp
    del col, dp, out


@pytest.mark.parametrize("kind", ["btree", "learned_plr", "binary_search", "hash"])
def test_size_sweep_plain_and_partitioned(gj, column, kind):
    # the hash table's three build paths: by 4096-slot range (2^24 keys), by 8192-slot range
    # (2^28 keys, a 2^29-slot table), by inserts (2^30 keys, 2^31 slots, 32 GiB)
    s, keys, col, dp, out, exp = column
    idx = gj.Index(kind, col, host_column=keys)
    st = idx.stats()
    print(f"[case] sweep 2^{s} {kind}: memory {st['memory_bytes'] / 2 ** 20:.0f} MiB layout={st['layout']}",
          flush=True)
    for approach, kw in (("plain", {}), ("partitioning", {"radix_bits": 10}),
                         ("partitioning", {"radix_bits": 12, "passes": 2})):
        print(f"[case] sweep 2^{s} {kind} {approach} {kw}", flush=True)
        out.fill(0x5A)
        idx.lookup(dp, out, approach=approach, **kw)
        gj.synchronize()
        assert np.array_equal(out.numpy(), exp), (approach, kw)
    assert idx.stats()["dropped_tuples"] == 0
    idx.close()


def test_harmonia_at_2_30(gj, column):
    s, keys, col, dp, out, exp = column
    if s != 30:
        pytest.skip("Harmonia is built at the largest size only")
    print("[case] sweep 2^30 harmonia", flush=True)
    idx = gj.Index("harmonia", col, host_column=keys)
    out.fill(0x5A)
    idx.lookup(dp, out)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)
    assert idx.stats()["depth"] >= 5
    idx.close()
