"""GPU parity for the radix partitioner and the lookup drivers (approaches).

The partitioner is compared with the oracle's stable partition (counts, the key
and payload sequence of every partition, 128-byte aligned starts, padding
markers).  Every approach must return the oracle's row ids for every index kind:
the output of the partitioned, blockwise-sorted and hash-join drivers is
independent of the partition layout.
"""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu

KINDS = ["binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash", "no_op"]


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("bits,passes", [(1, 1), (4, 1), (8, 1), (10, 1), (12, 1), (8, 2), (12, 2), (13, 2), (14, 2),
                                         (16, 2)])
@pytest.mark.parametrize("m", [0, 1, 1000, 100003])
def test_partition_matches_oracle(gj, orc, gpu, kb, bits, passes, m):
    print(f"[case] partition kb={kb} bits={bits} passes={passes} m={m}", flush=True)
    rng = np.random.default_rng(bits * 7 + m)
    dt = np.uint64 if kb == 8 else np.uint32
    keys = rng.integers(0, np.iinfo(dt).max, size=m, dtype=dt)
    ignore = 8 * kb - bits - 3
    d = gj.DeviceArray.from_numpy(keys) if m else gj.DeviceArray(1, dt)
    dropped = gj.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    dk, dp, off, cnt = gj.partition(d, m, bits, ignore, passes, dropped=dropped)
    gj.synchronize()
    assert int(dropped.numpy()[0]) == 0
    counts, ok, op = orc.partition(keys, bits, ignore)
    got_cnt = cnt.numpy()
    assert np.array_equal(got_cnt, counts)
    if m == 0:
        return
    offs = off.numpy()
    align = 128 // kb
    assert np.all(offs % align == 0)
    assert np.all(offs[1:] >= offs[:-1] + counts[:-1])
    gk, gp = dk.numpy(), dp.numpy()
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    for p in np.nonzero(counts)[0]:
        o, c, s = int(offs[p]), int(counts[p]), int(starts[p])
        assert np.array_equal(gk[o:o + c], ok[s:s + c]), p
        assert np.array_equal(gp[o:o + c], op[s:s + c]), p
        pad_end = o + (c + align - 1) // align * align
        assert np.all(gp[o + c:pad_end] == 0xFFFFFFFF)


@pytest.mark.parametrize("bits,passes", [(6, 1), (7, 1), (10, 1), (12, 1), (12, 2), (14, 2), (16, 2)])
def test_partition_many_tiles_per_block(gj, orc, gpu, bits, passes):
    # 5e6 keys: every block of the scatter walks several tiles (write-combining
    # scatter up to 7 bits, register tiles above), and the last tile of the last
    # block is partial
    m = 5_000_003
    print(f"[case] partition many tiles bits={bits} passes={passes} m={m}", flush=True)
    rng = np.random.default_rng(bits + 17 * passes)
    keys = rng.integers(0, 2 ** 64 - 1, size=m, dtype=np.uint64)
    ignore = 64 - bits
    dropped = gj.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    dk, dp, off, cnt = gj.partition(gj.DeviceArray.from_numpy(keys), m, bits, ignore, passes, dropped=dropped)
    gj.synchronize()
    assert int(dropped.numpy()[0]) == 0, "scatter dropped tuples"
    counts, ok, op = orc.partition(keys, bits, ignore)
    assert np.array_equal(cnt.numpy(), counts)
    offs = off.numpy().astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    # gather every partition's run from the padded layout into the unpadded order
    idx = np.repeat(offs - starts, counts.astype(np.int64)) + np.arange(m, dtype=np.int64)
    gk, gp = dk.numpy(), dp.numpy()
    assert np.array_equal(gk[idx], ok)
    assert np.array_equal(gp[idx], op)


@pytest.mark.parametrize("bits", [14, 16])
def test_partition_skewed_segments(gj, orc, gpu, bits):
    # 70 % of the keys share the top 8 bits, so the first pass leaves one segment with
    # most of the chunks (855 of them, not a multiple of the column scan's batch)
    m = 5_000_003
    print(f"[case] partition skewed bits={bits} m={m}", flush=True)
    rng = np.random.default_rng(bits + 0x5E)
    hot = int(m * 0.7)
    keys = np.concatenate([rng.integers(0, 2 ** 56, size=hot, dtype=np.uint64),
                           rng.integers(0, 2 ** 64 - 1, size=m - hot, dtype=np.uint64)])
    rng.shuffle(keys)
    ignore = 64 - bits
    dropped = gj.DeviceArray.from_numpy(np.zeros(1, np.uint64))
    dk, dp, off, cnt = gj.partition(gj.DeviceArray.from_numpy(keys), m, bits, ignore, 2, dropped=dropped)
    gj.synchronize()
    assert int(dropped.numpy()[0]) == 0, "scatter dropped tuples"
    counts, ok, op = orc.partition(keys, bits, ignore)
    assert np.array_equal(cnt.numpy(), counts)
    offs = off.numpy().astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    idx = np.repeat(offs - starts, counts.astype(np.int64)) + np.arange(m, dtype=np.int64)
    assert np.array_equal(dk.numpy()[idx], ok)
    assert np.array_equal(dp.numpy()[idx], op)


def test_partition_rejects_bad_args(gj, gpu):
    keys = gj.DeviceArray(16, np.uint64)
    with pytest.raises(gj.GjError):
        gj.partition(keys, 16, 17, 0)
    with pytest.raises(gj.GjError):
        gj.partition(keys, 16, 4, 64)


@pytest.mark.parametrize("kind", ["btree", "learned_plr", "hash"])
@pytest.mark.parametrize("pdist", [_lib.PROBES_UNIFORM, _lib.PROBES_ZIPF])
@pytest.mark.parametrize("kw,m", [
    ({"radix_bits": 10, "window": 1 << 12}, 300007),  # 74 windows in one batch, the last one ragged
    ({"radix_bits": 12, "window": 1 << 10}, 300007),  # 64 windows per batch (the partition cap), 5 batches
    ({"radix_bits": 12, "passes": 2, "window": 5000}, 300007),
    ({"radix_bits": 14, "passes": 2, "window": 30000}, 300007),  # 16 windows per batch, two segmented passes
    ({"radix_bits": 16, "window": 1 << 16}, 300007),  # 4 windows per batch
    # 33 windows, 4 write-combining chunks of 8 tiles each per window; 8448 partitions
    # (the multi-block starts scan)
    ({"radix_bits": 8, "window": 1 << 16}, 2_100_007),
    ({"radix_bits": 4, "window": 1}, 40003),  # one probe per window, 16384 windows per batch
    ({"radix_bits": 8, "window": 3}, 40003),
])
def test_partitioned_window_batches(gj, orc, gpu, kind, pdist, kw, m):
    # windows below 2^22 probes are partitioned a batch at a time, each window on its own
    print(f"[case] window batches {kind} probes={pdist} {kw} m={m}", flush=True)
    n = 1 << 18
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 0, 0x29)
    col = gj.DeviceArray.from_numpy(keys)
    params = gj.index_params(learned_layout=_lib.LAYOUT_LINES) if kind == "learned_plr" else None
    idx = gj.Index(kind, col, host_column=keys, params=params)
    probes, exp = gj.gen_probes(keys, m, pdist, 1.0, 0.1 if pdist == _lib.PROBES_UNIFORM else 0.0, 0, 0x2A)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(m, np.uint32)
    out.fill(0x5A)
    idx.lookup(dp, out, approach="partitioning", **kw)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)
    assert np.array_equal(orc.lookup_lower_bound(keys, probes), exp)
    assert idx.stats()["dropped_tuples"] == 0
    idx.close()


def _setup(gj, kind, n, kb, dist=_lib.KEYS_UNIFORM):
    n = (n // 32) * 32  # harmonia needs a multiple of 32
    keys = gj.gen_build_keys(n, kb, dist, 0, 0x21 + kb)
    col = gj.DeviceArray.from_numpy(keys)
    params = gj.index_params(learned_leaf_bits=12, mlp_steps=100)
    return keys, col, gj.Index(kind, col, host_column=keys, params=params)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("approach,extra", [
    ("bws", {}),
    ("partitioning", {"radix_bits": 8, "window": 1 << 14}),
    ("partitioning", {"radix_bits": 12, "passes": 2, "window": 100000}),
    ("partitioning", {"radix_bits": 10}),
    ("partitioning", {"radix_bits": 14, "window": 200000}),
    ("partitioning", {"radix_bits": 16}),
    ("hj", {}),
])
def test_approaches_match_oracle(gj, orc, gpu, kind, kb, approach, extra):
    print(f"[case] {kind} kb={kb} {approach} {extra}", flush=True)
    keys, col, idx = _setup(gj, kind, 1 << 18, kb)
    probes, _ = gj.gen_probes(keys, 300007, _lib.PROBES_UNIFORM, 1.0, 0.1, 0, 0x22)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    out.fill(0x5A)
    idx.lookup(dp, out, approach=approach, **extra)
    gj.synchronize()
    assert idx.stats()["dropped_tuples"] == 0
    want = orc.lookup_lower_bound(keys, probes)
    if kind == "no_op" and approach != "hj":
        want = np.full_like(want, gj.NOT_FOUND)
    assert np.array_equal(out.numpy(), want)


@pytest.mark.parametrize("kb,bits,passes", [(8, 8, 1), (8, 10, 1), (4, 7, 1), (4, 10, 1), (4, 11, 1), (4, 12, 2),
                                            (8, 12, 2), (8, 14, 2), (8, 16, 2), (4, 16, 2)])
def test_partitioned_lookup_many_tiles(gj, orc, gpu, kb, bits, passes):
    # 3e6 probes in one window: each block of the unordered write-combining scatter
    # walks several tiles, loading the next one ahead, and carries open segments between
    # them (u32 keys at fan-out 2048 use 4-slot segments).  Fan-outs whose buffers leave
    # room for two blocks per CU (u64 up to 2^8, u32 up to 2^7 per pass) take the
    # register-capped kernel, the others the uncapped one.
    print(f"[case] partitioned lookup many tiles kb={kb} bits={bits} passes={passes}", flush=True)
    keys, col, idx = _setup(gj, "btree", 1 << 20, kb)
    probes, _ = gj.gen_probes(keys, 3_000_017, _lib.PROBES_UNIFORM, 1.0, 0.05, 0, 0x24 + bits)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    out.fill(0x5A)
    idx.lookup(dp, out, approach="partitioning", radix_bits=bits, passes=passes, window=1 << 22)
    gj.synchronize()
    assert idx.stats()["dropped_tuples"] == 0
    assert np.array_equal(out.numpy(), orc.lookup_lower_bound(keys, probes))


@pytest.mark.parametrize("kw", [{"radix_bits": 10, "window": 1 << 18}, {"radix_bits": 16, "passes": 2}])
def test_partitioned_zipf_probes(gj, orc, gpu, kw):
    keys, col, idx = _setup(gj, "btree", 1 << 20, 8)
    probes, exp = gj.gen_probes(keys, 1 << 20, _lib.PROBES_ZIPF, 1.0, 0.0, 0, 0x23)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, approach="partitioning", **kw)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)
    assert np.array_equal(orc.lookup_lower_bound(keys, probes), exp)
