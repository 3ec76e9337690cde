"""gj_join_aggregate_noalloc (no reference counterpart): the allocation-free variant
of gj_join_aggregate, whose workgroup partial sums live in a static device array.
Its sums must equal gj_join_aggregate's (and the oracle's) exactly, on two calls in
a row, the second with a smaller fact batch, plus a batch small enough for the
direct path that skips the array."""
import ctypes as C

import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu


def _call(gj, fn, idx, cols, flag, lo, hi, result, rows, strategy=_lib.AGG_LANE_REFILL):
    fb = _lib.FactBatch(cols[0].ptr, cols[1].ptr, cols[2].ptr, cols[3].ptr, rows)
    ap = _lib.AggParams(lo, hi, flag.ptr, strategy, 0)
    _lib.check(_lib.call(fn, idx.handle, C.byref(fb), C.byref(ap), result.ptr, None))


@pytest.mark.parametrize("kind", ["learned_plr", "btree", "hash"])
def test_noalloc_join_aggregate_matches_allocating_one(gj, orc, gpu, kind):
    dim_rows = 200_000
    key, flag = gj.gen_dimension(dim_rows, 0x51)
    skey, sflag = gj.sort_dimension(key, flag)
    dkey = gj.DeviceArray.from_numpy(skey)
    dflag = gj.DeviceArray.from_numpy(sflag)
    idx = gj.Index(kind, dkey, host_column=skey)
    lo, hi = 100, 400  # a wide window, so many rows qualify
    # 3e6 rows need more than 128 workgroups (the static array is used); 4e5 rows is a
    # smaller input after it; 5e4 rows take the direct path that skips the array
    for rows, seed in ((3_000_000, 0x52), (400_000, 0x53), (50_000, 0x54)):
        print(f"[case] noalloc kind={kind} rows={rows}", flush=True)
        cols_h = gj.gen_fact(rows, dim_rows, 2557, 90000, 10500000, 10, seed)
        cols = [gj.DeviceArray.from_numpy(a) for a in cols_h]
        want = orc.join_aggregate(skey, sflag, *cols_h, lo, hi)
        for strategy in (_lib.AGG_LANE_REFILL, _lib.AGG_PLAIN, _lib.AGG_SORTED):
            res_a = gj.DeviceArray(4, np.int64)
            res_b = gj.DeviceArray(4, np.int64)
            res_a.fill(0)
            res_b.fill(0)
            _call(gj, "gj_join_aggregate", idx, cols, dflag, lo, hi, res_a, rows, strategy)
            _call(gj, "gj_join_aggregate_noalloc", idx, cols, dflag, lo, hi, res_b, rows, strategy)
            gj.synchronize()
            a = tuple(int(v) for v in res_a.numpy()[:3])
            b = tuple(int(v) for v in res_b.numpy()[:3])
            assert a == tuple(want), strategy
            assert b == a, strategy
    idx.close()


def test_noalloc_rejects_bad_arguments(gj, gpu):
    lib = _lib.load()
    assert lib.gj_join_aggregate_noalloc(None, None, None, None, None) == _lib.GJ_ERR_INVALID
    # the partitioned and lookup-first strategies need scratch memory: refused, nothing launched
    key, flag = gj.gen_dimension(1000, 0x55)
    skey, sflag = gj.sort_dimension(key, flag)
    idx = gj.Index("btree", gj.DeviceArray.from_numpy(skey), host_column=skey)
    cols = [gj.DeviceArray.from_numpy(a) for a in gj.gen_fact(100, 1000, 2557, 90000, 10500000, 10, 0x56)]
    res = gj.DeviceArray(4, np.int64)
    for strategy in (_lib.AGG_PARTITIONED, _lib.AGG_LOOKUP_FIRST):
        with pytest.raises(gj.GjError) as e:
            _call(gj, "gj_join_aggregate_noalloc", idx, cols, gj.DeviceArray.from_numpy(sflag), 0, 100, res, 100,
                  strategy)
        assert e.value.status == _lib.GJ_ERR_UNSUPPORTED
