"""GPU parity for the fact-dimension join-aggregate, batched plans and the autotuner.

Integer aggregates are exact: num, den and the match count equal the oracle's
int64 sums (the Q14 shape of src/tpch/tpch_14_hj.cu:38-58), for every index
kind over the dimension keys.  Plans run uneven batches by per-batch launches,
a replayed hipGraph and a persistent kernel; every replay is checked.
"""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu

AGG_KINDS = ["binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash"]


def _dimension(gj, rows, seed=0x41):
    key, flag = gj.gen_dimension(rows, seed)
    skey, sflag = gj.sort_dimension(key, flag)
    return skey, sflag


def _fact(gj, rows, dim_rows, seed=0x42):
    return gj.gen_fact(rows, dim_rows, 2557, 90000, 10500000, 10, seed)


def _agg_index(gj, kind, skey):
    col = gj.DeviceArray.from_numpy(skey)
    return col, gj.Index(kind, col, host_column=skey, params=gj.index_params(learned_leaf_bits=10, mlp_steps=100))


@pytest.mark.parametrize("strategy", ["lane_refill", "plain", "partitioned", "sorted", "lookup_first"])
@pytest.mark.parametrize("kind", AGG_KINDS)
def test_join_aggregate_matches_oracle(gj, orc, gpu, kind, strategy):
    print(f"[case] join-aggregate {kind} {strategy}", flush=True)
    dim_rows = 32 * 1000
    skey, sflag = _dimension(gj, dim_rows)
    # a quarter of the fact keys miss the dimension (keys above dim_rows)
    fkey, date, price, disc = _fact(gj, 2_000_003, dim_rows + dim_rows // 3)
    lo, hi = 1000, 1030
    col, idx = _agg_index(gj, kind, skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray(4, np.int64)
    res.fill(0)
    gj.join_aggregate(idx, *d, dflag, lo, hi, res, strategy=strategy)
    gj.synchronize()
    num, den, mt = (int(v) for v in res.numpy()[:3])
    onum, oden, omt = orc.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi)
    assert (num, den, mt) == (onum, oden, omt)
    assert gj.promo_revenue(num, den) == orc.promo_revenue(onum, oden)
    if strategy == "partitioned":
        assert idx.stats()["dropped_tuples"] == 0


@pytest.mark.parametrize("rows,lo,hi", [(1, 0, 2557), (16384, 0, 2557), (16385, 0, 400), (1_000_003, 0, 2557),
                                        (300_000, 3000, 3100)])
def test_join_aggregate_lookup_first_chunks(gj, orc, gpu, rows, lo, hi):
    # one or several 16384-row chunks, every row qualifying, some, or none (the result
    # then stays as it was)
    print(f"[case] lookup-first join-aggregate rows={rows} window=[{lo}, {hi})", flush=True)
    dim_rows = 32 * 900
    skey, sflag = _dimension(gj, dim_rows, seed=0x4D)
    fkey, date, price, disc = _fact(gj, rows, dim_rows + dim_rows // 5, seed=0x4E + rows)
    col, idx = _agg_index(gj, "btree", skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray.from_numpy(np.array([5, 7, 11, 0], dtype=np.int64))
    gj.join_aggregate(idx, *d, dflag, lo, hi, res, strategy="lookup_first")
    gj.synchronize()
    onum, oden, omt = orc.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi)
    assert tuple(int(v) for v in res.numpy()[:3]) == (5 + onum, 7 + oden, 11 + omt)


@pytest.mark.parametrize("bits", [1, 4, 11, 12])
@pytest.mark.parametrize("rows,lo,hi", [(1, 0, 2557), (5000, 0, 400), (700_001, 1000, 1030), (300_000, 3000, 3100)])
def test_join_aggregate_partitioned_fanouts(gj, orc, gpu, bits, rows, lo, hi):
    # fan-outs from 2 to 4096 over the fkey; (3000, 3100) is past the last date, so no
    # row qualifies and the result stays as it was
    print(f"[case] partitioned join-aggregate bits={bits} rows={rows} window=[{lo}, {hi})", flush=True)
    dim_rows = 32 * 700
    skey, sflag = _dimension(gj, dim_rows, seed=0x47)
    fkey, date, price, disc = _fact(gj, rows, dim_rows + dim_rows // 5, seed=0x48 + rows)
    col, idx = _agg_index(gj, "learned_plr", skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray.from_numpy(np.array([5, 7, 11, 0], dtype=np.int64))
    gj.join_aggregate(idx, *d, dflag, lo, hi, res, strategy="partitioned", radix_bits=bits)
    gj.synchronize()
    onum, oden, omt = orc.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi)
    assert tuple(int(v) for v in res.numpy()[:3]) == (5 + onum, 7 + oden, 11 + omt)  # added to the result
    assert idx.stats()["dropped_tuples"] == 0


def test_join_aggregate_rejects_bad_strategy(gj, gpu):
    skey, sflag = _dimension(gj, 1000, seed=0x49)
    col, idx = _agg_index(gj, "btree", skey)
    d = [gj.DeviceArray.from_numpy(a) for a in _fact(gj, 100, 1000, seed=0x4A)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray(4, np.int64)
    with pytest.raises(gj.GjError):
        gj.join_aggregate(idx, *d, dflag, 0, 10, res, strategy=7)
    with pytest.raises(gj.GjError):
        gj.join_aggregate(idx, *d, dflag, 0, 10, res, strategy="partitioned", radix_bits=17)


@pytest.mark.parametrize("rows", [1, 1024, 5000, 128 * 1024, 128 * 1024 + 1, 140 * 1024])
def test_join_aggregate_small_grids(gj, orc, gpu, rows):
    # up to 128 workgroups (1024 rows per workgroup step) add their sums to the
    # result themselves;
    # larger grids go through the partial buffer and the sum kernel
    dim_rows = 32 * 500
    skey, sflag = _dimension(gj, dim_rows, seed=0x45)
    fkey, date, price, disc = _fact(gj, rows, dim_rows + dim_rows // 4, seed=0x46 + rows)
    lo, hi = 0, 400  # about one row in six qualifies
    col, idx = _agg_index(gj, "btree", skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray(4, np.int64)
    strategies = ("lane_refill", "plain", "partitioned", "sorted", "lookup_first")
    for rep, strategy in enumerate(strategies):  # every call adds to the earlier calls' result
        if rep == 0:
            res.fill(0)
        gj.join_aggregate(idx, *d, dflag, lo, hi, res, strategy=strategy)
    gj.synchronize()
    onum, oden, omt = orc.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi)
    k = len(strategies)
    assert tuple(int(v) for v in res.numpy()[:3]) == (k * onum, k * oden, k * omt)


@pytest.mark.parametrize("kind", ["btree", "learned_plr", "radix_spline"])
@pytest.mark.parametrize("lo,hi", [(0, 2 ** 31 - 1), (0, 1300), (2400, 2500)])
def test_join_aggregate_sorted_refills(gj, orc, gpu, kind, lo, hi):
    # every row, about half of them, and a few percent qualify: the sorted strategy's
    # workgroups refill and sort their 2048-slot buffers many times, once, or never fill
    print(f"[case] sorted join-aggregate {kind} window=[{lo}, {hi})", flush=True)
    dim_rows = 32 * 3000
    skey, sflag = _dimension(gj, dim_rows, seed=0x4B)
    fkey, date, price, disc = _fact(gj, 3_000_017, dim_rows + dim_rows // 7, seed=0x4C)
    col, idx = _agg_index(gj, kind, skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    res = gj.DeviceArray(4, np.int64)
    res.fill(0)
    gj.join_aggregate(idx, *d, dflag, lo, hi, res, strategy="sorted")
    gj.synchronize()
    want = orc.join_aggregate(skey, sflag, fkey, date, price, disc, lo, hi)
    assert tuple(int(v) for v in res.numpy()[:3]) == tuple(want)
    assert want[2] > 0


def test_join_aggregate_all_rows_closed_form(gj, gpu):
    # every fact row qualifies and every dimension row is flagged: num == den == sum(price*(100-disc))
    dim_rows = 4096
    skey = np.arange(1, dim_rows + 1, dtype=np.uint32)
    sflag = np.ones(dim_rows, dtype=np.uint8)
    fkey, date, price, disc = _fact(gj, 100000, dim_rows, seed=9)
    col, idx = _agg_index(gj, "btree", skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    res = gj.DeviceArray(4, np.int64)
    res.fill(0)
    gj.join_aggregate(idx, *d, gj.DeviceArray.from_numpy(sflag), -1, 1 << 30, res)
    gj.synchronize()
    want = int(np.sum(price.astype(np.int64) * (100 - disc.astype(np.int64))))
    num, den, mt = (int(v) for v in res.numpy()[:3])
    assert num == den == want and mt == fkey.size


def _log_uniform_sizes(rng, total_max, lo=1e3, hi=1e6):
    sizes = []
    while sum(sizes) < total_max:
        sizes.append(int(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    return sizes


@pytest.mark.parametrize("mode", [_lib.BATCH_LAUNCH, _lib.BATCH_GRAPH, _lib.BATCH_PERSISTENT])
@pytest.mark.parametrize("kind", ["learned_plr", "hash", "btree"])
def test_aggregate_plan_batches(gj, orc, gpu, mode, kind):
    print(f"[case] aggregate plan mode={mode} kind={kind}", flush=True)
    dim_rows = 32 * 2000
    skey, sflag = _dimension(gj, dim_rows)
    rng = np.random.default_rng(mode)
    sizes = _log_uniform_sizes(rng, 3_000_000)
    rows = sum(sizes)
    fkey, date, price, disc = _fact(gj, rows, dim_rows, seed=0x43)
    col, idx = _agg_index(gj, kind, skey)
    d = [gj.DeviceArray.from_numpy(a) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag)
    lo, hi = 500, 530
    batches, off = [], 0
    for sz in sizes:
        batches.append((d[0].offset(off), d[1].offset(off), d[2].offset(off), d[3].offset(off), sz))
        off += sz
    res = gj.DeviceArray(4 * len(batches), np.int64)
    plan = gj.AggregatePlan(idx, batches, dflag, lo, hi, res, mode)
    want, off = [], 0
    for sz in sizes:
        s = slice(off, off + sz)
        want.append(orc.join_aggregate(skey, sflag, fkey[s], date[s], price[s], disc[s], lo, hi))
        off += sz
    want = np.array(want, dtype=np.int64)
    for rep in range(3):
        res.fill(0x5A)  # the plan zeroes its results before any batch adds to them
        gj.synchronize()
        plan.run()
        gj.synchronize()
        got = res.numpy().reshape(-1, 4)[:, :3]
        assert np.array_equal(got, want), f"replay {rep}"
        # eager work between replays
        eager = gj.DeviceArray(4, np.int64)
        eager.fill(0)
        gj.join_aggregate(idx, *d, dflag, lo, hi, eager, rows=sizes[0])
        gj.synchronize()
        assert np.array_equal(eager.numpy()[:3], want[0])


@pytest.mark.parametrize("mode", [_lib.BATCH_LAUNCH, _lib.BATCH_GRAPH, _lib.BATCH_PERSISTENT])
@pytest.mark.parametrize("kind", ["btree", "learned_plr", "harmonia"])
@pytest.mark.parametrize("kb", [4, 8])
def test_lookup_plan_batches(gj, orc, gpu, mode, kind, kb):
    print(f"[case] lookup plan mode={mode} kind={kind} kb={kb}", flush=True)
    keys = gj.gen_build_keys(32 * 8192, kb, _lib.KEYS_UNIFORM, 63 if kb == 8 else 0, 0x51)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(learned_leaf_bits=10))
    rng = np.random.default_rng(7 + mode)
    sizes = _log_uniform_sizes(rng, 2_000_000) + [0, 1]
    m = sum(sizes)
    probes, _ = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.1, 63 if kb == 8 else 0, 0x52)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(m, np.uint32)
    batches, off = [], 0
    for sz in sizes:
        batches.append((dp.offset(off), out.offset(off), sz))
        off += sz
    plan = gj.LookupPlan(idx, batches, mode)
    want = orc.lookup_lower_bound(keys, probes)
    for rep in range(3):
        out.fill(0x33)
        plan.run()
        gj.synchronize()
        assert np.array_equal(out.numpy(), want), f"replay {rep}"


def test_autotune_returns_exact_config(gj, orc, gpu):
    keys = gj.gen_build_keys(1 << 20, 8, _lib.KEYS_UNIFORM, 63, 0x61)
    col = gj.DeviceArray.from_numpy(keys)
    idx = gj.Index("btree", col, host_column=keys)
    probes, exp = gj.gen_probes(keys, 1 << 20, _lib.PROBES_UNIFORM, 1.0, 0.0, 63, 0x62)
    dp = gj.DeviceArray.from_numpy(probes)
    scratch = gj.DeviceArray(probes.size, np.uint32)
    best, rate = gj.autotune(idx, dp, scratch)
    print(f"[autotune] approach={best.approach} block={best.block_size} bits={best.radix_bits} "
          f"window={best.window} passes={best.passes} rate={rate / 1e9:.2f} G/s", flush=True)
    assert rate > 0
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out, params=best)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)


@pytest.mark.parametrize("kind,kb,field,allowed", [
    ("harmonia", 4, "harmonia_group_lanes", {8, 16, 32}),
    ("harmonia", 8, "harmonia_group_lanes", {8, 16}),
    ("btree", 8, "btree_lds_bytes", {4096, 16384, 32768}),
    ("radix_spline", 8, "rs_max_error", {16, 32, 64}),
    ("radix_spline", 8, "rs_radix_bits", {12, 13, 16, 18}),
    ("learned_plr", 8, "learned_leaf_bits", {0, 15, 16, 17}),
])
def test_index_autotuner_picks_a_candidate(gj, orc, gpu, kind, kb, field, allowed):
    # the index autotuner builds each candidate over the column, times it, and the
    # chosen parameters build an index that answers exactly
    print(f"[case] index autotuner {kind} kb={kb}", flush=True)
    keys = gj.gen_build_keys(1 << 20, kb, _lib.KEYS_UNIFORM, 0, 0x63 + kb)
    col = gj.DeviceArray.from_numpy(keys)
    probes, exp = gj.gen_probes(keys, 1 << 19, _lib.PROBES_UNIFORM, 1.0, 0.0, 0, 0x64)
    dp = gj.DeviceArray.from_numpy(probes)
    scratch = gj.DeviceArray(probes.size, np.uint32)
    best, rate = gj.autotune_index(kind, col, dp, scratch, host_column=keys)
    print(f"[autotune_index] {kind}: {field}={getattr(best, field)} rate={rate / 1e9:.2f} G/s", flush=True)
    assert rate > 0 and getattr(best, field) in allowed
    idx = gj.Index(kind, col, host_column=keys, params=best)
    out = gj.DeviceArray(probes.size, np.uint32)
    idx.lookup(dp, out)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)
    assert np.array_equal(orc.lookup_lower_bound(keys, probes), exp)
    idx.close()


@pytest.mark.parametrize("mode", [_lib.BATCH_LAUNCH, _lib.BATCH_GRAPH])
def test_plans_on_a_created_stream(gj, orc, gpu, mode):
    # launch and graph plans fork their batches over side streams from the caller's
    # stream and join them back: a run on a created stream, followed by host reads after
    # that stream's synchronisation, sees every batch's result
    print(f"[case] plans on a created stream mode={mode}", flush=True)
    stream = gj.Stream()
    keys = gj.gen_build_keys(32 * 4096, 8, _lib.KEYS_UNIFORM, 63, 0x71)
    col = gj.DeviceArray.from_numpy(keys, stream)
    idx = gj.Index("learned_plr", col, host_column=keys, stream=stream)
    sizes = _log_uniform_sizes(np.random.default_rng(0x72), 1_000_000)
    m = sum(sizes)
    probes, _ = gj.gen_probes(keys, m, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x73)
    dp = gj.DeviceArray.from_numpy(probes, stream)
    out = gj.DeviceArray(m, np.uint32)
    batches, off = [], 0
    for sz in sizes:
        batches.append((dp.offset(off), out.offset(off), sz))
        off += sz
    plan = gj.LookupPlan(idx, batches, mode, stream=stream)
    want = orc.lookup_lower_bound(keys, probes)
    for rep in range(2):
        out.fill(0x33, stream)
        plan.run(stream)
        stream.synchronize()
        assert np.array_equal(out.numpy(), want), f"lookup run {rep}"
    plan.close()

    dim_rows = 32 * 1000
    skey, sflag = _dimension(gj, dim_rows, seed=0x74)
    fsizes = _log_uniform_sizes(np.random.default_rng(0x75), 1_500_000)
    fkey, date, price, disc = _fact(gj, sum(fsizes), dim_rows, seed=0x76)
    acol, aidx = _agg_index(gj, "btree", skey)
    d = [gj.DeviceArray.from_numpy(a, stream) for a in (fkey, date, price, disc)]
    dflag = gj.DeviceArray.from_numpy(sflag, stream)
    fb, off = [], 0
    for sz in fsizes:
        fb.append((d[0].offset(off), d[1].offset(off), d[2].offset(off), d[3].offset(off), sz))
        off += sz
    res = gj.DeviceArray(4 * len(fb), np.int64)
    aplan = gj.AggregatePlan(aidx, fb, dflag, 0, 400, res, mode, stream=stream)
    want, off = [], 0
    for sz in fsizes:
        s = slice(off, off + sz)
        want.append(orc.join_aggregate(skey, sflag, fkey[s], date[s], price[s], disc[s], 0, 400))
        off += sz
    want = np.array(want, dtype=np.int64)
    for rep in range(2):
        res.fill(0x5A, stream)
        aplan.run(stream)
        stream.synchronize()
        assert np.array_equal(res.numpy().reshape(-1, 4)[:, :3], want), f"aggregate run {rep}"
    aplan.close()
