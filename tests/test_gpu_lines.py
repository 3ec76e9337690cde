"""GPU parity for the learned index's lines layout (GJ_LAYOUT_LINES).

The index keeps its own copy of the keys in 128-byte lines placed by the model;
every row id must still equal the oracle's equality-checked lower_bound over the
sorted column (hits, misses, smallest/largest keys), for both roots, every key
distribution, both key widths, every driver and the batch plans.  The placement
statistics of the root-only model are pinned by an independent restatement of
the placement in exact integer arithmetic.
"""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu

LINES = _lib.LAYOUT_LINES


def _probes(gj, keys, m, seed, miss=0.1):
    p, _ = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, miss, 0, seed)
    kmax = np.iinfo(keys.dtype).max
    edge = np.array([keys[0], keys[-1], 0, kmax - 1, max(int(keys[0]) - 1, 0), min(int(keys[-1]) + 1, kmax - 1)],
                    dtype=keys.dtype)
    return np.concatenate([edge, p])


def _lookup(gj, kind, keys, probes, params, **lp):
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=params)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    out.fill(0x5A)
    idx.lookup(dp, out, **lp)
    gj.synchronize()
    res = out.numpy()
    st = idx.stats()
    idx.close()
    return res, st


def _check_stats(st, n, kb):
    assert st["layout"] == LINES
    assert st["lines"] >= -(-n // (15 if kb == 8 else 30))
    assert 1.0 <= st["avg_lines"] <= st["max_lines"]
    assert st["fallback_keys"] == 0
    assert st["memory_bytes"] >= st["lines"] * 128


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
@pytest.mark.parametrize("dist", [_lib.KEYS_UNIFORM, _lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
@pytest.mark.parametrize("kb", [4, 8])
def test_lines_matches_oracle(gj, orc, gpu, kind, dist, kb):
    n = 1 << 18
    keys = gj.gen_build_keys(n, kb, dist, 0, 0x700 + dist)
    probes = _probes(gj, keys, 40000, 0x710 + dist)
    got, st = _lookup(gj, kind, keys, probes, gj.index_params(learned_layout=LINES, mlp_steps=150))
    print(f"[case] lines {kind} dist={dist} kb={kb}: lines={st['lines']} leaves={st['leaves']} "
          f"avg_lines={st['avg_lines']:.4f} max_lines={st['max_lines']}", flush=True)
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    _check_stats(st, n, kb)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("line_keys", [1, 7, 15])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 65539])
def test_lines_small_columns_and_loads(gj, orc, gpu, kb, line_keys, n):
    # line_keys = 15 fills the 8-byte lines completely, so nearly every key spills
    keys = np.unique(np.random.default_rng(n + kb).integers(1, 2 ** (8 * kb - 2), size=n,
                                                            dtype=np.uint64)).astype(
        np.uint64 if kb == 8 else np.uint32)
    probes = np.concatenate([keys, keys + 1, keys - 1, np.array([0, 1, np.iinfo(keys.dtype).max], keys.dtype)])
    for kind in ("learned_plr", "learned_mlp"):
        got, st = _lookup(gj, kind, keys, probes,
                          gj.index_params(learned_layout=LINES, learned_line_keys=line_keys, mlp_steps=20))
        assert np.array_equal(got, orc.lookup_lower_bound(keys, probes)), kind
        _check_stats(st, keys.size, kb)


def test_lines_dense_keys_rowid_equals_key(gj, gpu):
    n = 300007
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_DENSE, 0, 1)
    probes = np.random.default_rng(3).integers(0, n + 1000, size=100000).astype(np.uint64)
    got, st = _lookup(gj, "learned_plr", keys, probes, gj.index_params(learned_layout=LINES))
    assert np.array_equal(got, np.where(probes < n, probes, gj.NOT_FOUND).astype(np.uint32))
    assert st["max_lines"] == 1  # dense keys fill every line to exactly the target load


def _umulhi(a: int, b: int) -> int:
    return (a * b) >> 64


def test_lines_root_only_placement_pinned(gj, orc, gpu):
    # uniform keys with the default leaf count: the root-only model (no leaf table) wins;
    # its targets are exact integer arithmetic, so the placement statistics can be
    # restated here and must agree exactly
    n, lk = 1 << 16, 10
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x720)
    probes = _probes(gj, keys, 50000, 0x721)
    got, st = _lookup(gj, "learned_plr", keys, probes, gj.index_params(learned_layout=LINES))
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    assert st["leaves"] == 0 and st["depth"] == 1
    kmin, kmax = int(keys[0]), int(keys[-1])
    targets = -(-n // lk)
    mult = min((targets << 64) // (kmax - kmin + 1), (1 << 64) - 1)
    t = np.array([_umulhi(int(k) - kmin, mult) for k in keys], dtype=np.int64)
    assert t[0] == 0 and t[-1] < targets and np.all(np.diff(t) >= 0)
    c = np.bincount(t, minlength=targets)
    carry, line_of_first, sum_line, placed_lines = 0, [], 0, 0
    line = np.empty(n, dtype=np.int64)
    pos = 0
    for l in range(targets):
        avail = carry + int(c[l])
        placed = min(15, avail)
        carry = avail - placed
        line[pos:pos + placed] = l
        pos += placed
        placed_lines += 1
    while carry:
        placed = min(15, carry)
        line[pos:pos + placed] = placed_lines
        pos += placed
        carry -= placed
        placed_lines += 1
    assert pos == n
    spill = line - t
    assert st["lines"] == placed_lines
    assert st["max_lines"] == int(spill.max()) + 1
    assert abs(st["avg_lines"] - (1.0 + spill.mean())) < 1e-9


def test_lines_two_stage_leaves(gj, orc, gpu):
    # an explicit leaf count skips the root-only model: the root routes to 2^8 leaves
    keys = gj.gen_build_keys(1 << 18, 8, _lib.KEYS_LOGNORMAL, 63, 0x730)
    probes = _probes(gj, keys, 40000, 0x731)
    got, st = _lookup(gj, "learned_plr", keys, probes, gj.index_params(learned_layout=LINES, learned_leaf_bits=8))
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    assert st["leaves"] == 1 << 8 and st["depth"] == 2


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
def test_lines_every_driver_and_plan(gj, orc, gpu, kind):
    rng = np.random.default_rng(0x740)
    keys = np.unique(rng.integers(0, 2 ** 62, size=(1 << 20) + 3000, dtype=np.uint64))[: (1 << 20) + 77]
    probes = np.concatenate([rng.choice(keys, 200000), rng.integers(0, 2 ** 62, size=20000, dtype=np.uint64)])
    want = orc.lookup_lower_bound(keys, probes)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_layout=LINES, mlp_steps=100))
    assert idx.stats()["layout"] == LINES
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for approach, kw in (("plain", {}), ("bws", {}), ("partitioning", {"radix_bits": 10, "window": 1 << 16}),
                         ("partitioning", {"radix_bits": 12, "passes": 2}), ("hj", {})):
        print(f"[case] lines {kind} {approach} {kw}", flush=True)
        out.fill(0x5A)
        idx.lookup(dp, out, approach=approach, **kw)
        gj.synchronize()
        assert np.array_equal(out.numpy(), want), (approach, kw)
    sizes = [1, 1000, 50000, 0, 100000, probes.size - 151001]
    for mode in (gj.BATCH_PERSISTENT, gj.BATCH_GRAPH, gj.BATCH_LAUNCH):
        batches, off = [], 0
        for sz in sizes:
            batches.append((dp.offset(off), out.offset(off), sz))
            off += sz
        plan = gj.LookupPlan(idx, batches, mode)
        for rep in range(2):
            out.fill(0x33)
            plan.run()
            gj.synchronize()
            assert np.array_equal(out.numpy(), want), (mode, rep)
        plan.close()
    idx.close()


def test_auto_layout_prefers_lines_on_uniform_keys(gj, gpu):
    # 2^22 u64 keys = a 32 MiB column, past the size auto keeps in the column layout
    keys = gj.gen_build_keys(1 << 22, 8, _lib.KEYS_UNIFORM, 63, 0x750)
    col = gj.DeviceArray.from_numpy(keys)
    st = gj.Index("learned_plr", col, host_column=keys).stats()
    assert st["layout"] == LINES and st["avg_lines"] < 1.05
    st_col = gj.Index("learned_plr", col, host_column=keys,
                      params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN)).stats()
    assert st_col["layout"] == _lib.LAYOUT_COLUMN
    small = gj.gen_build_keys(1 << 20, 8, _lib.KEYS_UNIFORM, 63, 0x751)
    st_small = gj.Index("learned_plr", gj.DeviceArray.from_numpy(small), host_column=small).stats()
    assert st_small["layout"] == _lib.LAYOUT_COLUMN  # an 8 MiB column stays in the column layout
