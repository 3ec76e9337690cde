"""GPU parity: batched lookups through the C ABI against the oracle.

Row ids must be bit-identical to the oracle's equality-checked lower_bound
(binary_search_index semantics, src/indexes.cuh:357-361) for every index kind,
including misses, the smallest and largest keys, and ragged sizes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ["binary_search", "btree"]


def _column(gj, n, kb, seed, dist=1):
    return gj.gen_build_keys(n, kb, dist, 0, seed)


def _probes(gj, keys, m, seed, miss=0.1):
    if m == 0:
        return np.empty(0, dtype=keys.dtype)
    p, _ = gj.gen_probes(keys, m, 0, 1.0, miss, 0, seed)
    kmax = np.iinfo(keys.dtype).max
    edge = np.array([keys[0], keys[-1], 0, kmax, max(int(keys[0]) - 1, 0), min(int(keys[-1]) + 1, kmax)],
                    dtype=keys.dtype)
    return np.concatenate([edge, p])


def _run(gj, kind, keys, probes, **lp):
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys)
    m = probes.size
    dp = gj.DeviceArray.from_numpy(probes) if m else gj.DeviceArray(1, keys.dtype)
    out = gj.DeviceArray(max(m, 1), np.uint32)
    idx.lookup(dp, out, m=m, **lp)
    gj.synchronize()
    return out.numpy(m), idx


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("n", [1, 2, 17, 1000, 65539, 1 << 20])
def test_lookup_matches_oracle(gj, orc, gpu, kind, kb, n):
    print(f"[case] kind={kind} kb={kb} n={n}", flush=True)
    keys = _column(gj, n, kb, seed=0x11 + n)
    probes = _probes(gj, keys, 20000, seed=0x12 + n)
    got, _ = _run(gj, kind, keys, probes)
    want = orc.lookup_lower_bound(keys, probes)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", KINDS + ["no_op"])
def test_empty_probe_batch(gj, gpu, kind):
    keys = _column(gj, 1000, 8, seed=5)
    got, _ = _run(gj, kind, keys, np.empty(0, dtype=np.uint64))
    assert got.size == 0


def test_no_op_returns_not_found(gj, gpu):
    keys = _column(gj, 1000, 8, seed=6)
    probes = keys[::3].copy()
    got, _ = _run(gj, "no_op", keys, probes)
    assert np.all(got == gj.NOT_FOUND)


@pytest.mark.parametrize("kind", KINDS)
def test_dense_keys_rowid_equals_key(gj, gpu, kind):
    # dense build keys (generate_dense_dataset, src/index_lookup_common.cuh:44-49): row id == key
    n = 300007
    keys = gj.gen_build_keys(n, 8, 0, 0, 1)
    probes = np.random.default_rng(3).integers(0, n + 1000, size=100000).astype(np.uint64)
    got, _ = _run(gj, kind, keys, probes)
    want = np.where(probes < n, probes, gj.NOT_FOUND).astype(np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", KINDS)
def test_block_sizes(gj, orc, gpu, kind):
    keys = _column(gj, 50000, 8, seed=9)
    probes = _probes(gj, keys, 30000, seed=10)
    want = orc.lookup_lower_bound(keys, probes)
    for block in (64, 128, 256):
        got, _ = _run(gj, kind, keys, probes, block_size=block)
        assert np.array_equal(got, want), block


def test_unsorted_column_rejected(gj, gpu):
    keys = np.array([5, 3, 9, 11], dtype=np.uint64)
    col = gj.DeviceArray.from_numpy(keys)
    with pytest.raises(gj.GjError) as e:
        gj.Index("btree", col)
    assert e.value.status == -1


def test_duplicates_binary_search_first_occurrence(gj, orc, gpu):
    keys = np.array([1, 2, 2, 2, 5, 5, 9], dtype=np.uint64)
    probes = np.array([2, 5, 9, 1, 3, 0, 10], dtype=np.uint64)
    got, _ = _run(gj, "binary_search", keys, probes)
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))


ALL_KINDS = ["binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash"]


@pytest.mark.parametrize("kind", ALL_KINDS)
@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("approach", ["plain", "bws", "partitioning"])
def test_extreme_keys(gj, orc, gpu, kind, kb, approach):
    # keys at both ends of the key range (0 and the largest key below the all-ones
    # value, which the hash table reserves as its empty marker), a ragged column size
    # (not a multiple of a 128-byte line) and probes at and next to every extreme
    dt = np.uint64 if kb == 8 else np.uint32
    top = int(np.iinfo(dt).max) - 1
    mid = np.unique(np.random.default_rng(kb).integers(2, top - 2, size=4000, dtype=dt))
    keys = np.unique(np.concatenate([np.array([0, 1, top - 1, top], dtype=dt), mid]))
    if kind == "harmonia":
        keys = keys[: keys.size // 32 * 32]
        keys[-1] = top
    probes = np.concatenate([keys, np.array([0, 1, 2, top - 2, top - 1, top, top + 1], dtype=dt),
                             keys[1:-1] + 1, keys[1:-1] - 1])
    print(f"[case] extreme keys kind={kind} kb={kb} approach={approach} n={keys.size}", flush=True)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_leaf_bits=6, mlp_steps=50))
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, approach=approach, radix_bits=6, window=1000)
    gj.synchronize()
    assert np.array_equal(out.numpy(), orc.lookup_lower_bound(keys, probes))


@pytest.mark.parametrize("kind", ["binary_search", "btree", "radix_spline", "learned_plr", "learned_mlp", "hash"])
@pytest.mark.parametrize("kb", [4, 8])
def test_single_key_column(gj, orc, gpu, kind, kb):
    dt = np.uint64 if kb == 8 else np.uint32
    keys = np.array([12345], dtype=dt)
    probes = np.array([0, 12344, 12345, 12346, np.iinfo(dt).max - 1], dtype=dt)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_leaf_bits=4, mlp_steps=20))
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out)
    gj.synchronize()
    assert np.array_equal(out.numpy(), orc.lookup_lower_bound(keys, probes))


@pytest.mark.parametrize("kind", ["binary_search", "btree", "radix_spline", "learned_plr", "hash"])
@pytest.mark.parametrize("approach", ["plain", "partitioning"])
def test_host_resident_column(gj, orc, gpu, kind, approach):
    # the column stays in page-locked host memory; the index's own structures and the
    # probes are in HBM, and the kernels read column lines over the host link
    keys = _column(gj, 100003, 8, seed=0x71)
    probes = _probes(gj, keys, 30000, seed=0x72)
    col = gj.HostArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, approach=approach, radix_bits=8, window=10000)
    gj.synchronize()
    assert np.array_equal(out.numpy(), orc.lookup_lower_bound(keys, probes))
    idx.close()
    col.free()
