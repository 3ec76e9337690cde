"""CPU tests that pin the oracle before it is trusted as the GPU path's checker.

Each check can fail for a plausible mistake in the restatement (an off-by-one in
a bound, a dropped level, a transposed operand): reference known answers, closed
forms, brute force on tiny inputs, invariants, and the reference's own Zipf
sampler compiled from the reference tree (oracle/_ref).
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ---- search primitives -----------------------------------------------------------

def test_lower_bound_kat_bf_search(orc):
    # src/bf_search.cpp:168-186: every needle 0..23 of this array against std::lower_bound
    data = np.array([1, 2, 4, 5, 5, 6, 8, 12, 20, 20, 21, 23], dtype=np.uint32)
    for needle in range(0, 24):
        assert orc.lower_bound(data, needle) == int(np.searchsorted(data, needle, side="left"))


def test_lower_bound_brute_force(orc):
    rng = np.random.default_rng(1)
    for n in range(0, 40):
        arr = np.sort(rng.integers(0, 20, size=n)).astype(np.uint64)
        for x in range(-1, 22):
            x = max(x, 0)
            want = next((i for i in range(n) if arr[i] >= x), n)
            assert orc.lower_bound(arr, x) == want


def test_branchy_binary_search(orc):
    rng = np.random.default_rng(2)
    for n in range(1, 60):
        arr = np.unique(rng.integers(0, 200, size=n)).astype(np.uint64)
        for x in range(0, 202):
            pos = orc.branchy_binary_search(arr, arr.size, x)
            assert pos == int(np.searchsorted(arr, x))  # unique keys: equal to lower_bound
    dup = np.array([1, 2, 2, 2, 3], dtype=np.uint64)
    assert dup[orc.branchy_binary_search(dup, 5, 2)] == 2  # some equal element


def test_lookup_lower_bound_semantics(orc):
    keys = np.array([3, 7, 7, 10], dtype=np.uint64)
    probes = np.array([3, 7, 10, 0, 4, 11, 2 ** 64 - 1], dtype=np.uint64)
    assert list(orc.lookup_lower_bound(keys, probes)) == [0, 1, 3] + [orc.NOT_FOUND] * 4
    assert orc.checksum(np.array([1, 2, 0xFFFFFFFF], dtype=np.uint32)) == 1 + 2 + 0xFFFFFFFF


# ---- B-tree with the reference page layout -----------------------------------------

def test_btree_page_capacities(orc):
    # LeafNode / InnerNode::maxEntries for 4 KiB pages with a 128 B header (src/btree.cuh:30-63)
    assert orc.btree_capacities(8, 4) == (329, 246)  # u64 keys, u32 row ids
    assert orc.btree_capacities(8, 8) == (247, 246)  # u64 keys, u64 row ids (index_lookup)
    assert orc.btree_capacities(4, 4) == (495, 329)


@pytest.mark.parametrize("n", [1, 2, 328, 329, 330, 329 * 246, 329 * 247, 329 * 247 + 1, 100003, 1 << 20])
def test_btree_equals_lower_bound(orc, n):
    rng = np.random.default_rng(n)
    keys = np.unique(rng.integers(0, 2 ** 63, size=n + n // 10 + 10, dtype=np.uint64))[:n]
    t = orc.BTree(keys, 4)
    probes = np.concatenate([keys, rng.integers(0, 2 ** 63, size=2000, dtype=np.uint64),
                             np.array([0, 2 ** 64 - 1], dtype=np.uint64)])
    assert np.array_equal(t.lookup(probes), orc.lookup_lower_bound(keys, probes))
    info = t.info()
    chunks = -(-n // 329)
    # one level of inner nodes whenever the chunks fit one node (clustered index, src/btree.cuh:177-181)
    assert (info["depth"] == 1) == (chunks <= 247)


# ---- Harmonia ------------------------------------------------------------------------

@pytest.mark.parametrize("n,fanout", [(280, 9), (2000, 9), (144, 5)])
def test_harmonia_reference_kats(orc, n, fanout):
    # test/test_harmonia.cu:37-55,115-123 and test/main.cu:144-164: lookup(keys[i]) == i over iota keys
    keys = np.arange(n, dtype=np.uint32)
    assert np.array_equal(orc.Harmonia(keys, fanout).lookup(keys), keys)


@pytest.mark.parametrize("n", [7, 15])
def test_harmonia_kats_small_sizes(orc, n):
    # test_root_only / host_lookup(15), fanout 9: with leaf keys aliasing the column the reference
    # rejects sizes that are not a multiple of 8 (src/harmonia.cuh:350-352); with copied leaf keys
    # the same trees answer every key
    keys = np.arange(n, dtype=np.uint32)
    with pytest.raises(ValueError):
        orc.Harmonia(keys, 9)
    h = orc.Harmonia(keys, 9, reuse_input=False)
    assert np.array_equal(h.lookup(keys), keys)
    assert h.depth() == (1 if n == 7 else 2)


def test_harmonia_structure_280_fanout_9(orc):
    # gather_tree_info / store_structure by hand: 35 leaves -> 4 nodes (8, 8, 8, 7 keys) -> root (3 keys)
    h = orc.Harmonia(np.arange(280, dtype=np.uint32) * 2, 9)
    keys, children = h.arrays()
    assert h.depth() == 3
    assert list(children[:5]) == [1, 5, 14, 23, 32]
    assert list(children[5:8]) == [0, 8, 16]
    # root separators: largest key under level-1 nodes 0..2 (each spans 9 leaves of 8 keys)
    assert list(keys[:3]) == [2 * (72 - 1), 2 * (144 - 1), 2 * (216 - 1)]
    assert keys[3] == 0xFFFFFFFF  # padding with the largest u32 key


def test_harmonia_random_equals_lower_bound(orc):
    rng = np.random.default_rng(4)
    for n in [32, 64, 32 * 33, 32 * 34, 32 * 33 * 33 + 32 * 5, 32 * 4000]:
        keys = np.unique(rng.integers(0, 2 ** 63, size=2 * n, dtype=np.uint64))[:n]
        probes = np.concatenate([keys, rng.integers(0, 2 ** 63, size=3000, dtype=np.uint64)])
        assert np.array_equal(orc.Harmonia(keys).lookup(probes), orc.lookup_lower_bound(keys, probes))


def test_harmonia_golden(orc):
    g = _gold("harmonia_kat.npz")
    for n, f in [(280, 9), (2000, 9), (144, 5), (32 * 34, 33), (32 * 1100, 33)]:
        keys = np.arange(n, dtype=np.uint32) * 3 + 1
        h = orc.Harmonia(keys, f)
        k, c = h.arrays()
        assert np.array_equal(k, g[f"keys_{n}_{f}"]) and np.array_equal(c, g[f"children_{n}_{f}"])
        assert h.depth() == int(g[f"depth_{n}_{f}"][0])


# ---- RadixSpline -----------------------------------------------------------------------

@pytest.mark.parametrize("err", [1, 8, 32])
@pytest.mark.parametrize("dist", [1, 2, 3])
def test_radix_spline_error_invariant(gj, orc, err, dist):
    keys = gj.gen_build_keys(30000, 8, dist, 0, 40 + dist)
    rs = orc.RadixSpline(keys, 12, err)
    # every build key is estimated within the bound, so the bounded search is exact
    assert rs.max_abs_error() <= err + 1e-6
    px, py, radix = rs.arrays()
    assert px[0] == keys[0] and px[-1] == keys[-1]
    assert np.all(px[1:] > px[:-1])
    pos = np.searchsorted(keys, px)
    assert np.array_equal(keys[pos], px) and np.array_equal(py, pos.astype(np.float64))
    assert np.all(np.diff(radix.astype(np.int64)) >= 0) and radix[-1] == px.size
    probes = np.concatenate([keys, np.random.default_rng(5).integers(0, 2 ** 64 - 1, size=3000, dtype=np.uint64)])
    assert np.array_equal(rs.lookup(probes), orc.lookup_lower_bound(keys, probes))


def test_radix_spline_golden(orc):
    g = _gold("rs_u64.npz")
    rs = orc.RadixSpline(g["keys"], 10, 8)
    px, py, radix = rs.arrays()
    assert np.array_equal(px, g["knot_x"]) and np.array_equal(py, g["knot_y"]) and np.array_equal(radix, g["radix"])


def test_radix_spline_estimate_endpoints(orc):
    keys = np.array([10, 20, 30, 40, 1000], dtype=np.uint64)
    rs = orc.RadixSpline(keys, 4, 1)
    assert rs.estimate(5) == 0.0 and rs.estimate(10) == 0.0  # key <= min
    assert rs.estimate(1000) == 4.0 and rs.estimate(5000) == 4.0  # key >= max


# ---- hash table, partitioner, join-aggregate ------------------------------------------------

def test_hash_capacity_closed_form(orc):
    # 2^ceil(log2(n / 0.5)), src/linear_probing_hashtable.cuh:74-79
    assert orc.hash_capacity(2 ** 26) == 2 ** 27
    assert orc.hash_capacity(1000) == 2048
    assert orc.hash_capacity(1024) == 2048
    assert orc.hash_capacity(1025) == 4096


def test_murmur3_finaliser(orc):
    assert orc.murmur3_64(0) == 0 and orc.murmur3_32(0) == 0  # fixed point of the xorshift-multiply chain
    # bijective on a sample (distinct inputs give distinct hashes)
    xs = np.arange(1, 20001, dtype=np.uint64)
    assert len({orc.murmur3_64(int(x)) for x in xs}) == xs.size


# The multiplicative inverses of MurmurHash3's finaliser constants (mod 2^64 for fmix64,
# mod 2^32 for fmix32), as published with the finaliser's inversion; an xorshift by s
# bits is undone by xoring the shifts by s, 2s, 3s, ...
_INV64 = (0x4F74430C22A54005, 0x9CB4B2F8129337DB)  # of 0xff51afd7ed558ccd, 0xc4ceb9fe1a85ec53
_INV32 = (0xA5CB9243, 0x7ED1B41D)  # of 0x85ebca6b, 0xc2b2ae35


def _unshift(h, s, bits):
    out, t = h, h >> s
    while t:
        out ^= t
        t >>= s
    return out & ((1 << bits) - 1)


def _unfmix(h, shifts, inv, bits):
    m = (1 << bits) - 1
    h = _unshift(h, shifts[2], bits)
    h = (h * inv[1]) & m
    h = _unshift(h, shifts[1], bits)
    h = (h * inv[0]) & m
    return _unshift(h, shifts[0], bits)


def test_murmur3_finaliser_inverts(orc):
    # undoing each step with independently published constants recovers the input, which a
    # mistyped multiplier or shift in the oracle's finalisers would break
    rng = np.random.default_rng(0x3E)
    for x in [0, 1, 2, 3, 0x123456789ABCDEF, 2 ** 63, 2 ** 64 - 1] + [int(v) for v in
                                                                      rng.integers(0, 2 ** 63, 500, dtype=np.uint64)]:
        assert _unfmix(orc.murmur3_64(x), (33, 33, 33), _INV64, 64) == x
        x32 = x & 0xFFFFFFFF
        assert _unfmix(orc.murmur3_32(x32), (16, 13, 16), _INV32, 32) == x32


def test_hash_join_equals_lower_bound(orc):
    rng = np.random.default_rng(6)
    for kb in (4, 8):
        dt = np.uint64 if kb == 8 else np.uint32
        keys = np.unique(rng.integers(0, np.iinfo(dt).max - 1, size=5000, dtype=dt))
        probes = np.concatenate([keys[::3], rng.integers(0, np.iinfo(dt).max - 1, size=3000, dtype=dt)])
        assert np.array_equal(orc.hash_join(keys, probes), orc.lookup_lower_bound(keys, probes))


def test_partition_brute_force(orc):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 2 ** 64 - 1, size=20000, dtype=np.uint64)
    counts, ok, op = orc.partition(keys, 6, 50)
    digit = (keys >> np.uint64(50)) & np.uint64(63)
    assert np.array_equal(counts, np.bincount(digit.astype(np.int64), minlength=64).astype(np.uint64))
    order = np.argsort(digit, kind="stable")
    assert np.array_equal(ok, keys[order]) and np.array_equal(op, order.astype(np.uint32))


def test_join_aggregate_brute_force(orc):
    g = _gold("join_agg.npz")
    lo, hi = (int(v) for v in g["window"])
    got = orc.join_aggregate(g["dim_key"], g["dim_flag"], g["fkey"], g["date"], g["price"], g["disc"], lo, hi)
    sel = (g["date"] >= lo) & (g["date"] < hi)
    pos = np.searchsorted(g["dim_key"], g["fkey"].astype(np.uint32))
    pos_c = np.minimum(pos, g["dim_key"].size - 1)
    hit = sel & (g["dim_key"][pos_c] == g["fkey"].astype(np.uint32))
    s = g["price"].astype(np.int64) * (100 - g["disc"].astype(np.int64))
    want = (int(s[hit & (g["dim_flag"][pos_c] != 0)].sum()), int(s[hit].sum()), int(hit.sum()))
    assert got == want == tuple(int(v) for v in g["sums"])


def test_promo_revenue_formula(orc):
    # process_results, src/tpch/tpch_14_common.cu:124-127
    assert orc.promo_revenue(0, 0) == 0
    assert orc.promo_revenue(16380778, 100000000) == 100 * 16380778 * 1000 // (100000000 // 1000)
    assert orc.promo_revenue(1, 999) == 0  # den / 1000 == 0


# ---- Zipf, RNG, MLP gradients ---------------------------------------------------------------

@pytest.mark.parametrize("n,q,seed", [(100, 1.0, 1), (2 ** 26 - 1, 1.0, 7), (1000, 0.5, 3), (1000, 1.25, 5),
                                      (50, 2.0, 11)])
def test_zipf_matches_reference_sampler(orc, n, q, seed):
    # the restated rejection-inversion sampler equals the reference's own zipf_distribution
    # (src/zipf.hpp), fed by the same std::mt19937 seed; fixtures hold the reference's draws
    g = _gold("zipf_ref.npz")
    mine = orc.zipf_mt19937(n, q, seed, 2000)
    assert np.array_equal(mine, g[f"draws_{n}_{q}_{seed}"])
    if orc.ref_zipf_available():
        assert np.array_equal(mine, orc.ref_zipf_mt19937(n, q, seed, 2000))
    assert mine.min() >= 1 and mine.max() <= n


def test_splitmix64_published_values(orc):
    # SplitMix64 started at state 0 (Steele, Lea, Flood 2014)
    assert orc.splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0, 1) == 0x6E789E6AA1B965F4


def test_mlp_gradient_finite_differences(orc):
    rng = np.random.default_rng(8)
    h = 8
    p = rng.normal(0, 1.0, size=3 * h + 1).astype(np.float32)
    p[:h] = np.abs(p[:h]) * 3
    x = rng.random(500).astype(np.float32)
    y = (x ** 0.5).astype(np.float32)
    loss, grad = orc.mlp_loss_grad(p, h, x, y)
    eps = 1e-3
    for i in range(3 * h + 1):
        pp, pm = p.astype(np.float64).copy(), p.astype(np.float64).copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = orc.mlp_loss_grad(pp.astype(np.float32), h, x, y)
        lm, _ = orc.mlp_loss_grad(pm.astype(np.float32), h, x, y)
        num = (lp - lm) / (np.float32(p[i] + eps) - np.float32(p[i] - eps))
        assert abs(num - grad[i]) <= 2e-2 * max(1e-3, abs(grad[i])) + 2e-4, i
    out = orc.mlp_forward(p, h, x)
    assert np.allclose(((out - y) ** 2).mean(), loss, rtol=1e-4)
