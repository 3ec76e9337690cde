"""CPU tests of the seeded workload generators (libgj host code)."""
import numpy as np
import pytest

from gpujoin_amd import _lib


def test_rng_is_splitmix64(gj, orc):
    for seed in (0, 1, 0x1234567890ABCDEF):
        for i in (0, 1, 77, 1 << 40):
            assert gj.rng_u64(seed, i) == orc.splitmix64(seed, i)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("dist", [_lib.KEYS_UNIFORM, _lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
def test_build_keys_sorted_unique_deterministic(gj, kb, dist):
    a = gj.gen_build_keys(50000, kb, dist, 0, 99, threads=1)
    b = gj.gen_build_keys(50000, kb, dist, 0, 99, threads=8)
    assert np.array_equal(a, b)  # independent of the thread count
    assert a.size == 50000 and np.all(a[1:] > a[:-1]) and np.unique(a).size == a.size
    c = gj.gen_build_keys(50000, kb, dist, 0, 100)
    assert not np.array_equal(a, c)


def test_uniform_keys_cover_range(gj):
    k = gj.gen_build_keys(1 << 16, 8, _lib.KEYS_UNIFORM, 63, 5)
    assert k.max() < 2 ** 63
    # uniform over [0, 2^63): the quartiles sit near 1/4, 1/2, 3/4 of the range
    q = np.quantile(k.astype(np.float64), [0.25, 0.5, 0.75]) / 2.0 ** 63
    assert np.allclose(q, [0.25, 0.5, 0.75], atol=0.01)
    d = gj.gen_build_keys(1000, 4, _lib.KEYS_DENSE, 0, 0)
    assert np.array_equal(d, np.arange(1000, dtype=np.uint32))


def test_u32_keys_from_full_range(gj):
    # 2^20 unique u32 keys drawn from [0, 2^32) (the CPU-runnable configuration)
    k = gj.gen_build_keys(1 << 20, 4, _lib.KEYS_UNIFORM, 0, 1)
    assert np.unique(k).size == 1 << 20 and k.max() > 0xF0000000


def test_probe_expected_row_ids(gj, orc):
    keys = gj.gen_build_keys(10000, 8, _lib.KEYS_UNIFORM, 63, 3)
    probes, exp = gj.gen_probes(keys, 50000, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 4)
    hit = exp != gj.NOT_FOUND
    assert np.array_equal(keys[exp[hit]], probes[hit])
    assert abs(hit.mean() - 0.9) < 0.01
    assert np.array_equal(orc.lookup_lower_bound(keys, probes), exp)  # misses are absent keys
    p2, e2 = gj.gen_probes(keys, 50000, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 4, threads=3)
    assert np.array_equal(p2, probes) and np.array_equal(e2, exp)


def test_zipf_probe_ranks(gj):
    keys = gj.gen_build_keys(100000, 8, _lib.KEYS_UNIFORM, 63, 3)
    probes, exp = gj.gen_probes(keys, 400000, _lib.PROBES_ZIPF, 1.0, 0.0, 63, 5)
    assert np.array_equal(keys[exp], probes)
    cnt = np.bincount(exp, minlength=4)
    # the reference never draws rank 0 (k in [1, n-1], src/index_lookup_common.cuh:147)
    assert cnt[0] == 0
    # Zipf theta = 1: P(rank 1) / P(rank 2) = 2
    assert abs(cnt[1] / cnt[2] - 2.0) < 0.15


def test_dimension_and_fact(gj):
    key, flag = gj.gen_dimension(200000, 7)
    assert np.array_equal(np.sort(key), np.arange(1, 200001, dtype=np.int32))
    assert abs(flag.mean() - 1 / 3) < 0.01
    sk, sf = gj.sort_dimension(key, flag)
    assert np.array_equal(sk, np.arange(1, 200001, dtype=np.uint32))
    assert np.array_equal(sf[key - 1], flag)
    fkey, date, price, disc = gj.gen_fact(300000, 200000, 2557, 90000, 10500000, 10, 8)
    assert fkey.min() >= 1 and fkey.max() <= 200000
    assert date.min() >= 0 and date.max() < 2557
    assert price.min() >= 90000 and price.max() <= 10500000
    assert disc.min() == 0 and disc.max() == 10
    # a 30-day window selects ~30/2557 of the rows
    sel = ((date >= 1000) & (date < 1030)).mean()
    assert abs(sel - 30 / 2557) < 0.002


@pytest.mark.parametrize("n,q", [(99_999, 1.0), ((1 << 26) - 1, 1.0), (1000, 0.5), (1_000_000, 1.5), (2, 1.0)])
def test_zipf_trial_equals_oracle_on_equal_uniforms(gj, orc, n, q):
    # libgj's rejection-inversion trial (what GJ_PROBES_ZIPF draws ranks with) against the
    # oracle's, which is pinned bit for bit to the reference's zipf_distribution (src/zipf.hpp)
    # by the committed draws of tests/golden/zipf_ref.npz: equal uniforms, equal ranks
    lib = gj._lib.load()
    rng = np.random.default_rng(n)
    us = np.concatenate([np.linspace(0.0, 1.0, 4001, endpoint=False), rng.random(4000),
                         [0.0, 1e-12, 0.5, 1.0 - 2.0 ** -53]])
    got = [int(lib.gj_zipf_trial(n, q, float(u))) for u in us]
    want = [orc.zipf_trial(n, q, float(u)) for u in us]
    assert got == want
    accepted = [k for k in got if k]
    assert accepted and min(accepted) >= 1 and max(accepted) <= n
