"""GPU parity for Harmonia, RadixSpline, the learned models and the hash table.

Every row id is compared bit for bit with the oracle's equality-checked
lower_bound on the same seeded inputs (hits, misses, smallest/largest keys).
"""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu


def _probes(gj, keys, m, seed, miss=0.1):
    p, _ = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, miss, 0, seed)
    kmax = np.iinfo(keys.dtype).max
    edge = np.array([keys[0], keys[-1], 0, kmax - 1, max(int(keys[0]) - 1, 0), min(int(keys[-1]) + 1, kmax - 1)],
                    dtype=keys.dtype)
    return np.concatenate([edge, p])


def _lookup(gj, kind, keys, probes, params=None, **lp):
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=params)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, **lp)
    gj.synchronize()
    res = out.numpy()
    st = idx.stats()
    idx.close()
    return res, st


# ---- Harmonia ---------------------------------------------------------------

@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("n", [32, 64, 32 * 33, 32 * 34, 32 * 33 * 33 + 32 * 7, 1 << 20])
def test_harmonia_matches_oracle(gj, orc, gpu, kb, n):
    print(f"[case] harmonia kb={kb} n={n}", flush=True)
    keys = gj.gen_build_keys(n, kb, _lib.KEYS_UNIFORM, 0, 100 + n)
    probes = _probes(gj, keys, 20000, 200 + n)
    want = orc.lookup_lower_bound(keys, probes)
    for g in ([8, 16] if kb == 8 else [8, 16, 32]):
        got, st = _lookup(gj, "harmonia", keys, probes, params=gj.index_params(harmonia_group_lanes=g))
        assert np.array_equal(got, want), f"group lanes {g}"
    # the reference Harmonia built over the same column has the same depth and answers
    h = orc.Harmonia(keys, 33)
    assert st["depth"] == h.depth()
    assert np.array_equal(h.lookup(probes), want)


def test_harmonia_requires_multiple_of_32(gj, gpu):
    keys = gj.gen_build_keys(1000, 8, _lib.KEYS_UNIFORM, 0, 7)
    col = gj.DeviceArray.from_numpy(keys)
    with pytest.raises(gj.GjError) as e:
        gj.Index("harmonia", col)
    assert e.value.status == _lib.GJ_ERR_INVALID


def test_harmonia_dense_rowid_equals_key(gj, gpu):
    n = 32 * 40000
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_DENSE, 0, 1)
    probes = np.random.default_rng(5).integers(0, n + 500, size=50000).astype(np.uint64)
    got, _ = _lookup(gj, "harmonia", keys, probes)
    assert np.array_equal(got, np.where(probes < n, probes, gj.NOT_FOUND).astype(np.uint32))


# ---- RadixSpline ------------------------------------------------------------

@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("err", [8, 32, 64])
@pytest.mark.parametrize("dist", [_lib.KEYS_UNIFORM, _lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
def test_radix_spline_matches_oracle(gj, orc, gpu, kb, err, dist):
    n = 200003
    print(f"[case] radix_spline kb={kb} err={err} dist={dist}", flush=True)
    keys = gj.gen_build_keys(n, kb, dist, 0, 300 + err + dist)
    probes = _probes(gj, keys, 30000, 400 + err)
    want = orc.lookup_lower_bound(keys, probes)
    got, st = _lookup(gj, "radix_spline", keys, probes, params=gj.index_params(rs_max_error=err, rs_radix_bits=18))
    assert np.array_equal(got, want)
    ref = orc.RadixSpline(keys, 18, err)
    px, _, _ = ref.arrays()
    assert st["spline_points"] == px.size  # same GreedySplineCorridor knot count as the oracle's builder
    assert ref.max_abs_error() <= err + 1
    assert np.array_equal(ref.lookup(probes), want)


@pytest.mark.parametrize("bits", [1, 4, 12, 13, 14, 18])
@pytest.mark.parametrize("lds_bits", [6, 13, 16])
def test_radix_spline_radix_bits(gj, orc, gpu, bits, lds_bits):
    # the radix table is staged in LDS when it has at most 2^lds_bits + 2 entries
    keys = gj.gen_build_keys(50000, 8, _lib.KEYS_UNIFORM, 0, 17)
    probes = _probes(gj, keys, 20000, 18)
    got, st = _lookup(gj, "radix_spline", keys, probes,
                      params=gj.index_params(rs_radix_bits=bits, rs_lds_bits=lds_bits, rs_max_error=16))
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    assert st["lds_levels"] == (1 if bits <= lds_bits else 0)


# ---- learned two-stage models -------------------------------------------------

@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
@pytest.mark.parametrize("dist", [_lib.KEYS_UNIFORM, _lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
@pytest.mark.parametrize("kb", [4, 8])
def test_learned_matches_oracle(gj, orc, gpu, kind, dist, kb):
    n = 1 << 18
    print(f"[case] {kind} dist={dist} kb={kb}", flush=True)
    keys = gj.gen_build_keys(n, kb, dist, 0, 500 + dist)
    probes = _probes(gj, keys, 40000, 600 + dist)
    want = orc.lookup_lower_bound(keys, probes)
    params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=10, learned_max_error=16, mlp_steps=150)
    got, st = _lookup(gj, kind, keys, probes, params=params)
    assert np.array_equal(got, want)
    assert st["leaves"] == 1 << 10
    assert st["max_window"] <= 2 * 16 + 2


def test_learned_fallback_is_exact(gj, orc, gpu):
    # a tiny bound forces most leaves onto the B-tree fallback; answers stay exact
    keys = gj.gen_build_keys(1 << 17, 8, _lib.KEYS_LOGNORMAL, 0, 77)
    probes = _probes(gj, keys, 30000, 78)
    got, st = _lookup(gj, "learned_plr", keys, probes,
                      params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=6, learned_max_error=1))
    assert st["fallback_leaves"] > 0
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
def test_learned_few_fallback_leaves_exact(gj, orc, gpu, kind):
    # uniform keys plus one run of 600 consecutive integers: only the leaves over the
    # run exceed the error bound, so fewer than 1 key in 400 falls back and the plain
    # driver marks those probes and re-looks them up with the B-tree in a second pass
    rng = np.random.default_rng(91)
    base = np.unique(rng.integers(0, 2 ** 62, size=1 << 20, dtype=np.uint64))
    run = np.arange(2 ** 61, 2 ** 61 + 600, dtype=np.uint64)
    keys = np.unique(np.concatenate([base, run]))
    probes = np.concatenate([_probes(gj, keys, 200000, 92), run, run + 600])
    params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=14, learned_max_error=16, learned_root_segments=1,
                             mlp_steps=100)
    got, st = _lookup(gj, kind, keys, probes, params=params)
    print(f"[case] {kind} fallback_keys={st['fallback_keys']} of {keys.size}", flush=True)
    assert 0 < st["fallback_keys"] and st["fallback_keys"] * 400 <= keys.size
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
@pytest.mark.parametrize("kb", [4, 8])
def test_split_leaves_exact(gj, orc, gpu, kind, kb):
    # clustered keys under a linear root: whole clusters land in single leaves, which are
    # split into 16 equal-count sub-leaves (one separator line picks the sub-leaf); the
    # sub-leaves that are still too wide go to the B-tree; every answer stays exact
    keys = gj.gen_build_keys(1 << 18, kb, _lib.KEYS_CLUSTERED, 0, 0x51 + kb)
    probes = _probes(gj, keys, 60000, 0x52 + kb)
    params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=12, learned_max_error=32, learned_root_segments=1,
                             mlp_steps=60)
    got, st = _lookup(gj, kind, keys, probes, params=params)
    print(f"[case] {kind} kb={kb} split_leaves={st['split_leaves']} fallback_keys={st['fallback_keys']}", flush=True)
    assert st["split_leaves"] > 0
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    # the same through every driver
    for approach in ("bws", "partitioning"):
        got2, _ = _lookup(gj, kind, keys, probes, params=params, approach=approach, radix_bits=8, window=20000)
        assert np.array_equal(got2, got), approach


def test_mlp_root_beats_linear_root_on_skewed_keys(gj, gpu):
    # what the trained network gains: on lognormal keys the linear root crowds most
    # keys into a few leaves; the trained monotone MLP root spreads them, so far fewer
    # keys need the B-tree fallback and the mean last-mile window shrinks
    keys = gj.gen_build_keys(1 << 20, 8, _lib.KEYS_LOGNORMAL, 63, 0x31)
    params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=12, learned_max_error=32)
    col = gj.DeviceArray.from_numpy(keys)
    linear = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=12, learned_max_error=32, learned_root_segments=1)
    plr = gj.Index("learned_plr", col, host_column=keys, params=linear).stats()
    assert plr["root_segments"] == 0
    mlp = gj.Index("learned_mlp", col, host_column=keys, params=params).stats()
    print(f"[gain] plr fallback_keys={plr['fallback_keys']} avg_window={plr['avg_window']:.1f}; "
          f"mlp fallback_keys={mlp['fallback_keys']} avg_window={mlp['avg_window']:.1f} loss={mlp['train_loss']:.2e}",
          flush=True)
    assert mlp["fallback_keys"] < plr["fallback_keys"]


@pytest.mark.parametrize("segments", [0, 2, 64, 2048, 4096])
@pytest.mark.parametrize("dist", [_lib.KEYS_UNIFORM, _lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
@pytest.mark.parametrize("kb", [4, 8])
def test_learned_spline_root_matches_oracle(gj, orc, gpu, segments, dist, kb):
    n = 1 << 18
    print(f"[case] spline root segments={segments} dist={dist} kb={kb}", flush=True)
    keys = gj.gen_build_keys(n, kb, dist, 0, 700 + dist)
    probes = _probes(gj, keys, 40000, 800 + dist)
    params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=10, learned_max_error=16,
                             learned_root_segments=segments)
    got, st = _lookup(gj, "learned_plr", keys, probes, params=params)
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    if segments >= 2:
        assert 1 <= st["root_segments"] <= segments


def test_learned_spline_root_tiny_columns(gj, orc, gpu):
    # fewer keys than requested segments: the root keeps at most n segments
    for n in (1, 2, 3, 5, 64, 100):
        keys = np.unique(np.random.default_rng(n).integers(1, 2 ** 40, size=n, dtype=np.uint64))
        probes = np.concatenate([keys, keys + 1, keys - 1])
        params = gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=4, learned_root_segments=4096)
        got, st = _lookup(gj, "learned_plr", keys, probes, params=params)
        assert np.array_equal(got, orc.lookup_lower_bound(keys, probes)), n
        assert st["root_segments"] <= max(2, keys.size)


def test_spline_root_spreads_clustered_keys(gj, gpu):
    # 64 narrow clusters: a linear root sends each cluster to one leaf, and even the
    # leaves' 16-way splits leave many keys to the B-tree; the automatic choice switches
    # to a spline of the column's CDF, which spreads the keys evenly over the leaves
    keys = gj.gen_build_keys(1 << 20, 8, _lib.KEYS_CLUSTERED, 63, 0x33)
    col = gj.DeviceArray.from_numpy(keys)
    lin = gj.Index("learned_plr", col, host_column=keys,
                   params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=12, learned_max_error=32,
                                         learned_root_segments=1)).stats()
    auto = gj.Index("learned_plr", col, host_column=keys,
                    params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=12, learned_max_error=32)).stats()
    print(f"[gain] linear root fallback_keys={lin['fallback_keys']}; auto root segments={auto['root_segments']} "
          f"fallback_keys={auto['fallback_keys']} avg_window={auto['avg_window']:.1f}", flush=True)
    assert lin["root_segments"] == 0 and lin["fallback_keys"] > 0 and lin["split_leaves"] > 0
    assert 1 <= auto["root_segments"] <= 2048
    assert auto["fallback_keys"] * 20 < lin["fallback_keys"]


def test_mlp_gradients_match_oracle(gj, orc, gpu):
    rng = np.random.default_rng(11)
    h = 32
    p = rng.normal(0, 0.5, size=3 * h + 1).astype(np.float32)
    x = rng.random(4096).astype(np.float32)
    y = np.sqrt(x).astype(np.float32)
    dp, dx, dy = gj.DeviceArray.from_numpy(p), gj.DeviceArray.from_numpy(x), gj.DeviceArray.from_numpy(y)
    dg = gj.DeviceArray(3 * h + 1, np.float32)
    dl = gj.DeviceArray(1, np.float32)
    gj.mlp_loss_grad(dp, h, dx, dy, dg, dl)
    gj.synchronize()
    loss_ref, grad_ref = orc.mlp_loss_grad(p, h, x, y)
    # FP32 accumulation on the GPU vs double on the CPU: 1e-4 relative to the gradient scale
    tol = 1e-4 * max(1.0, float(np.abs(grad_ref).max()))
    assert np.allclose(dg.numpy(), grad_ref, rtol=1e-3, atol=tol)
    assert abs(float(dl.numpy()[0]) - loss_ref) <= 1e-4 * max(1.0, loss_ref)
    out = gj.DeviceArray(x.size, np.float32)
    gj.mlp_forward(dp, h, dx, out)
    gj.synchronize()
    assert np.allclose(out.numpy(), orc.mlp_forward(p, h, x), rtol=1e-5, atol=1e-5)


# ---- hash table ---------------------------------------------------------------

@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_hash_matches_oracle(gj, orc, gpu, kb, n):
    keys = gj.gen_build_keys(n, kb, _lib.KEYS_UNIFORM, 0, 900 + n)
    probes = _probes(gj, keys, 30000, 901 + n)
    want = orc.lookup_lower_bound(keys, probes)
    got, st = _lookup(gj, "hash", keys, probes)
    assert np.array_equal(got, want)
    assert np.array_equal(orc.hash_join(keys, probes), want)
    assert st["memory_bytes"] >= orc.hash_capacity(n) * 16


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
@pytest.mark.parametrize("kb", [4, 8])
def test_compact_leaves_exact(gj, orc, gpu, kind, kb):
    # from 2^22 keys, the default leaf count gives twice the leaves in the 8-byte form
    # (group bases in LDS) when every leaf stays within the bound; row ids stay exact,
    # misses included
    rng = np.random.default_rng(93 + kb)
    dt = np.uint64 if kb == 8 else np.uint32
    hi = 2 ** 62 if kb == 8 else 2 ** 32 - 2
    keys = np.unique(rng.integers(0, hi, size=(1 << 22) + 40000, dtype=dt))[: (1 << 22) + 12345]
    probes = np.concatenate([rng.choice(keys, 300000), rng.integers(0, hi, size=50000, dtype=dt),
                             keys[:100], keys[-100:]]).astype(dt)
    got, st = _lookup(gj, kind, keys, probes, params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, mlp_steps=100))
    print(f"[case] {kind} kb={kb} leaves={st['leaves']} avg_window={st['avg_window']:.1f}", flush=True)
    assert st["leaves"] == 1 << 17 and st["fallback_leaves"] == 0
    assert np.array_equal(got, orc.lookup_lower_bound(keys, probes))
    # an explicit leaf count keeps the 16-byte leaves and the same answers
    got16, st16 = _lookup(gj, kind, keys, probes,
                          params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, learned_leaf_bits=16, mlp_steps=100))
    assert st16["leaves"] == 1 << 16
    assert np.array_equal(got16, got)


@pytest.mark.parametrize("kind", ["learned_plr", "learned_mlp"])
def test_compact_leaves_every_driver(gj, orc, gpu, kind):
    # the compact-leaf descriptor through the blockwise-sorting, partitioned and
    # hash-join drivers and the persistent and graph lookup plans
    rng = np.random.default_rng(95)
    keys = np.unique(rng.integers(0, 2 ** 62, size=(1 << 22) + 30000, dtype=np.uint64))[: (1 << 22) + 777]
    probes = np.concatenate([rng.choice(keys, 400000), rng.integers(0, 2 ** 62, size=40000, dtype=np.uint64)])
    want = orc.lookup_lower_bound(keys, probes)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_layout=_lib.LAYOUT_COLUMN, mlp_steps=100))
    assert idx.stats()["leaves"] == 1 << 17
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for approach, kw in (("bws", {}), ("partitioning", {"radix_bits": 10, "window": 1 << 17}),
                         ("partitioning", {"radix_bits": 12, "passes": 2}), ("hj", {})):
        print(f"[case] {kind} compact {approach} {kw}", flush=True)
        out.fill(0x5A)
        idx.lookup(dp, out, approach=approach, **kw)
        gj.synchronize()
        assert np.array_equal(out.numpy(), want), (approach, kw)
    sizes = [1, 1000, 50000, 0, 250000, probes.size - 301001]
    for mode in (gj.BATCH_PERSISTENT, gj.BATCH_GRAPH):
        print(f"[case] {kind} compact plan mode={mode}", flush=True)
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


def _rate(gj, idx, dp, out, m, steps=3):
    s = gj.Stream()
    idx.lookup(dp, out, m=m, stream=s)
    s.synchronize()
    e0, e1 = gj.Event(), gj.Event()
    e0.record(s)
    for _ in range(steps):
        idx.lookup(dp, out, m=m, stream=s)
    e1.record(s)
    s.synchronize()
    return m * steps / (e0.elapsed_ms(e1) * 1e-3)


@pytest.mark.parametrize("dist", [_lib.KEYS_LOGNORMAL, _lib.KEYS_CLUSTERED])
def test_what_training_gains(gj, orc, gpu, dist):
    # The MLP root with its GPU training against the same root left at its initialisation
    # (the piecewise-linear CDF through 33 quantiles), in both layouts, on 2^22 skewed keys
    # with 10 % misses.  Training must lower the root's loss.  What that buys, measured:
    #  - column layout, lognormal keys: the trained root spreads the keys so that no leaf
    #    needs the B-tree (76 078 fallback keys untrained), about 2x the lookup rate;
    #  - clustered keys: no gain: 64 narrow clusters defeat a 32-ramp root either way (every
    #    key falls back in the column layout; the lines layout's leaves absorb it).
    # The spline root of learned_plr is printed alongside as the method without a network.
    n, m = 1 << 22, 1 << 23
    keys = gj.gen_build_keys(n, 8, dist, 63, 0x31 + dist)
    probes, exp = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x32 + dist)
    col, dp = gj.DeviceArray.from_numpy(keys), gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(m, np.uint32)
    rows = {}
    for layout in (_lib.LAYOUT_COLUMN, _lib.LAYOUT_LINES):
        for label, flags in (("trained", 0), ("untrained", _lib.LEARNED_UNTRAINED)):
            idx = gj.Index("learned_mlp", col, host_column=keys,
                           params=gj.index_params(learned_layout=layout, learned_flags=flags))
            st = idx.stats()
            rate = _rate(gj, idx, dp, out, m)
            assert np.array_equal(out.numpy(), exp), (layout, label)
            rows[(layout, label)] = (st, rate)
            print(f"[gain] dist={dist} layout={layout} {label}: loss={st['train_loss']:.3e} "
                  f"avg_window={st['avg_window']:.1f} fallback_keys={st['fallback_keys']} "
                  f"avg_lines={st['avg_lines']:.4f} max_lines={st['max_lines']} trained_root={st['root_trained']} "
                  f"{rate / 1e9:.2f} G lookups/s", flush=True)
            idx.close()
        assert rows[(layout, "trained")][0]["train_loss"] < rows[(layout, "untrained")][0]["train_loss"]
        assert rows[(layout, "untrained")][0]["root_trained"] == 0
    spline = gj.Index("learned_plr", col, host_column=keys)
    print(f"[gain] dist={dist} learned_plr (spline root) layout={spline.stats()['layout']}: "
          f"{_rate(gj, spline, dp, out, m) / 1e9:.2f} G lookups/s", flush=True)
    spline.close()
    if dist == _lib.KEYS_LOGNORMAL:
        tr, un = rows[(_lib.LAYOUT_COLUMN, "trained")], rows[(_lib.LAYOUT_COLUMN, "untrained")]
        assert tr[0]["fallback_keys"] < un[0]["fallback_keys"]
        assert tr[1] > 1.3 * un[1]


def _btree_levels(n, nk):
    """Inner levels of the implicit B-tree: the chunk count divided by nk until one node."""
    c, levels = -(-n // nk), 0
    while True:
        c = -(-c // nk)
        levels += 1
        if c <= 1:
            return levels


@pytest.mark.parametrize("nb", [64, 128, 256])
@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("n", [1, 7, 100, 4097, 300007])
def test_btree_node_sizes(gj, orc, gpu, nb, kb, n):
    # nodes (and leaf chunks) of half a line, one line and two lines, with and without
    # LDS-resident upper levels, through the plain, bws and partitioned drivers
    print(f"[case] btree node={nb} kb={kb} n={n}", flush=True)
    dt = np.uint64 if kb == 8 else np.uint32
    rng = np.random.default_rng(nb + kb * 7 + n)
    keys = np.unique(rng.integers(1, np.iinfo(dt).max - 1, size=n + n // 4 + 8, dtype=dt))[:n]
    probes = np.concatenate([keys, keys + 1, keys - 1, rng.integers(0, np.iinfo(dt).max, size=2000, dtype=dt),
                             np.array([0, np.iinfo(dt).max], dtype=dt)])
    want = orc.lookup_lower_bound(keys, probes)
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for lds in (1, 16384):
        idx = gj.Index("btree", col, host_column=keys, params=gj.index_params(btree_node_bytes=nb, btree_lds_bytes=lds))
        st = idx.stats()
        assert st["depth"] == _btree_levels(keys.size, nb // kb), st
        assert (st["lds_levels"] == 0) == (lds == 1)
        for approach, kw in (("plain", {}), ("bws", {}), ("partitioning", {"radix_bits": 8, "window": 3000})):
            out.fill(0x5A)
            idx.lookup(dp, out, approach=approach, **kw)
            gj.synchronize()
            assert np.array_equal(out.numpy(), want), (lds, approach)
        idx.close()


def test_btree_node_size_rejected(gj, gpu):
    keys = np.arange(1, 1001, dtype=np.uint64)
    with pytest.raises(gj.GjError):
        gj.Index("btree", gj.DeviceArray.from_numpy(keys), host_column=keys, params=gj.index_params(btree_node_bytes=96))


@pytest.mark.parametrize("kb", [4, 8])
def test_hash_large_column(gj, orc, gpu, kb):
    # a 2^21-key table (4 M slots of 16 bytes): the hash index and the hash-join driver,
    # which rebuilds the table on every call
    n = (1 << 21) + 12345
    dt = np.uint64 if kb == 8 else np.uint32
    rng = np.random.default_rng(0x4A5 + kb)
    keys = np.unique(rng.integers(1, np.iinfo(dt).max - 1, size=n + n // 8, dtype=dt))[:n]
    probes = np.concatenate([rng.choice(keys, 300000), rng.integers(0, np.iinfo(dt).max, size=30000, dtype=dt),
                             keys[:3], keys[-3:]])
    want = orc.lookup_lower_bound(keys, probes)
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx = gj.Index("hash", col, host_column=keys)
    out.fill(0x5A)
    idx.lookup(dp, out)
    gj.synchronize()
    assert np.array_equal(out.numpy(), want)
    assert idx.stats()["dropped_tuples"] == 0
    bt = gj.Index("btree", col, host_column=keys)
    for rep in range(2):  # the driver rebuilds the table on every call
        out.fill(0x5A)
        bt.lookup(dp, out, approach="hj")
        gj.synchronize()
        assert np.array_equal(out.numpy(), want), rep
    assert bt.stats()["dropped_tuples"] == 0


def _unfmix(h, kb):
    """Keys whose murmur3 finaliser (of the key's width) is h: each step of the finaliser
    undone in reverse order, with the multipliers' inverses mod 2^64 / 2^32."""
    bits = 8 * kb
    dt = np.uint64 if kb == 8 else np.uint32
    shifts, inv = ((33, 33, 33), (0x4F74430C22A54005, 0x9CB4B2F8129337DB)) if kb == 8 else \
        ((16, 13, 16), (0xA5CB9243, 0x7ED1B41D))

    def unshift(x, s):
        out, t = x.copy(), x >> dt(s)
        while np.any(t):
            out ^= t
            t >>= dt(s)
        return out

    x = h.astype(dt)
    with np.errstate(over="ignore"):
        x = unshift(x, shifts[2])
        x = x * dt(inv[1])
        x = unshift(x, shifts[1])
        x = x * dt(inv[0])
        x = unshift(x, shifts[0])
    assert bits in (32, 64)
    return x



@pytest.mark.parametrize("kb", [4, 8])
def test_hash_crowded_ranges(gj, orc, gpu, kb):
    # keys chosen by their hashes: 12000 home slots in the first 8192 slots of the 2^17-slot
    # table (its first two 4096-slot build ranges) and 3000 in its last 200 slots, so the
    # build meets ranges with more keys than slots (their excess runs on through the next
    # ranges) and runs that wrap from the last slot to the first; 30000 more spread evenly
    cap = 1 << 17
    dt = np.uint64 if kb == 8 else np.uint32
    rng = np.random.default_rng(0x7A + kb)

    def keys_with_homes(lo, hi, count):
        high = rng.integers(0, np.iinfo(dt).max, size=count, dtype=dt) & ~dt(cap - 1)
        return _unfmix(high | rng.integers(lo, hi, size=count).astype(dt), kb)

    keys = np.unique(np.concatenate([keys_with_homes(0, 8192, 12000), keys_with_homes(cap - 200, cap, 3000),
                                     keys_with_homes(0, cap, 30000)]))
    keys = keys[keys != np.iinfo(np.uint64).max]  # the empty-slot marker
    fin = orc.murmur3_64 if kb == 8 else orc.murmur3_32
    homes = np.array([fin(int(k)) & (cap - 1) for k in keys])
    assert np.count_nonzero(homes < 8192) > 11900 and np.count_nonzero(homes >= cap - 200) > 2900
    assert (1 << 16) < 2 * keys.size <= cap  # the table the build sizes has 2^17 slots
    probes = np.concatenate([keys, rng.integers(0, np.iinfo(dt).max, size=20000, dtype=dt)])
    want = orc.lookup_lower_bound(keys, probes)
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx = gj.Index("hash", col, host_column=keys)
    out.fill(0x5A)
    idx.lookup(dp, out)
    gj.synchronize()
    assert np.array_equal(out.numpy(), want)
    bs = gj.Index("binary_search", col, host_column=keys)
    out.fill(0x5A)
    bs.lookup(dp, out, approach="hj")
    gj.synchronize()
    assert np.array_equal(out.numpy(), want)
    assert idx.stats()["dropped_tuples"] == 0 and bs.stats()["dropped_tuples"] == 0


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("n", [1000, 1008, 4096 + 3])
def test_btree_key_max_in_last_chunk(gj, orc, gpu, kb, n):
    # the column ends in the largest key, and its last leaf chunk is partial (n % 16 != 0
    # for u64), so the chunk's tail holds the same value as fill: a probe of that key must
    # find the real position, and the fill past the end must never answer
    dt = np.uint64 if kb == 8 else np.uint32
    kmax = np.iinfo(dt).max
    rng = np.random.default_rng(0x99 + n + kb)
    keys = np.unique(rng.integers(0, kmax - 1, size=n + n // 4, dtype=dt))[:n - 1]
    keys = np.concatenate([keys, np.array([kmax], dtype=dt)])
    probes = np.concatenate([np.array([kmax, kmax - 1, keys[-2], 0], dtype=dt), rng.choice(keys, 500)])
    want = orc.lookup_lower_bound(keys, probes)
    assert want[0] == keys.size - 1
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for nb in (64, 128, 256):
        idx = gj.Index("btree", col, host_column=keys, params=gj.index_params(btree_node_bytes=nb))
        out.fill(0x5A)
        idx.lookup(dp, out)
        gj.synchronize()
        assert np.array_equal(out.numpy(), want), nb
        idx.close()
