"""GPU parity with the index structures in page-locked host memory (GJ_PLACE_HOST),
the reference's default placement (src/index_lookup_config.hpp:52-66: indexes in
host memory, read over the interconnect; migrated pages src/btree.cuh:440-473).
Row ids must equal the oracle's for every kind, through the plain and partitioned
drivers, with the column in HBM or in host memory as well."""
import numpy as np
import pytest

from gpujoin_amd import _lib

pytestmark = pytest.mark.gpu

KINDS = ["binary_search", "btree", "harmonia", "radix_spline", "learned_plr", "learned_mlp", "hash"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("column_in_host", [False, True])
def test_index_in_host_memory(gj, orc, gpu, kind, column_in_host):
    print(f"[case] placement host kind={kind} column_in_host={column_in_host}", flush=True)
    n = (1 << 18) // 32 * 32
    keys = gj.gen_build_keys(n, 8, _lib.KEYS_UNIFORM, 63, 0x81)
    probes, exp = gj.gen_probes(keys, 100000, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x82)
    col = gj.HostArray.from_numpy(keys) if column_in_host else gj.DeviceArray.from_numpy(keys)
    params = gj.index_params(placement=_lib.PLACE_HOST, learned_layout=_lib.LAYOUT_LINES, mlp_steps=60)
    idx = gj.Index(kind, col, host_column=keys, params=params)
    st = idx.stats()
    assert st["placement"] == _lib.PLACE_HOST
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for approach, kw in (("plain", {}), ("partitioning", {"radix_bits": 10})):
        out.fill(0x5A)
        idx.lookup(dp, out, approach=approach, **kw)
        gj.synchronize()
        assert np.array_equal(out.numpy(), exp), approach
    idx.close()
    if column_in_host:
        col.free()


def test_auto_layout_in_host_memory_is_column(gj, orc, gpu):
    # a 16 MiB column, where auto picks the lines layout in HBM, keeps the column layout
    # when the index structures go to host memory
    keys = gj.gen_build_keys(1 << 21, 8, _lib.KEYS_UNIFORM, 63, 0x83)
    probes, exp = gj.gen_probes(keys, 200000, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x84)
    col = gj.DeviceArray.from_numpy(keys)
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    for placement, layout in ((_lib.PLACE_DEVICE, _lib.LAYOUT_LINES), (_lib.PLACE_HOST, _lib.LAYOUT_COLUMN)):
        idx = gj.Index("learned_plr", col, host_column=keys, params=gj.index_params(placement=placement))
        assert idx.stats()["layout"] == layout, placement
        out.fill(0x5A)
        idx.lookup(dp, out)
        gj.synchronize()
        assert np.array_equal(out.numpy(), exp), placement
        idx.close()


@pytest.mark.parametrize("kind", ["btree", "learned_plr"])
@pytest.mark.parametrize("column_in_host", [False, True])
def test_autotuned_lookup_in_host_memory(gj, orc, gpu, kind, column_in_host):
    # the lookup autotuner over an index whose structures sit in host memory (the
    # configuration the bench's host-placement leg measures): the configuration it
    # returns must give the oracle's row ids
    print(f"[case] autotune host placement kind={kind} column_in_host={column_in_host}", flush=True)
    keys = gj.gen_build_keys(1 << 20, 8, _lib.KEYS_UNIFORM, 63, 0x86)
    probes, exp = gj.gen_probes(keys, 1 << 20, _lib.PROBES_UNIFORM, 1.0, 0.1, 63, 0x87)
    col = gj.HostArray.from_numpy(keys) if column_in_host else gj.DeviceArray.from_numpy(keys)
    idx = gj.Index(kind, col, host_column=keys, params=gj.index_params(placement=_lib.PLACE_HOST))
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(probes.size, np.uint32)
    best, rate = gj.autotune(idx, dp, out)
    print(f"[autotune host] approach={best.approach} bits={best.radix_bits} passes={best.passes} "
          f"window={best.window} rate={rate / 1e9:.2f} G/s", flush=True)
    assert rate > 0
    out.fill(0x5A)
    idx.lookup(dp, out, params=best)
    gj.synchronize()
    assert np.array_equal(out.numpy(), exp)
    assert idx.stats()["dropped_tuples"] == 0
    idx.close()
    if column_in_host:
        col.free()


@pytest.mark.parametrize("kind", ["btree", "learned_plr", "harmonia"])
def test_large_column_in_host_memory(gj, gpu, kind):
    # an 8 GiB dense column (2^30 keys, past 2^32-byte offsets) in page-locked host memory
    # with the index structures there too, as in the reference's host-memory sweep
    # (scripts/bench_tlb_misses.sh); dense keys, so key k sits at row k, and misses are
    # the keys past the column's end
    n = 1 << 30
    print(f"[case] host column 2^30 dense kind={kind}", flush=True)
    hcol = gj.HostArray(n, np.uint64)
    view = hcol.view()
    chunk = 1 << 26
    for c in range(0, n, chunk):
        view[c:c + chunk] = np.arange(c, c + chunk, dtype=np.uint64)
    rng = np.random.default_rng(0x85)
    m = 1 << 22
    probes = rng.integers(0, n + n // 8, size=m, dtype=np.uint64)
    probes[:4] = [0, n - 1, n, 2 ** 63]
    exp = np.where(probes < n, probes, gj.NOT_FOUND).astype(np.uint32)
    idx = gj.Index(kind, hcol, host_column=view, params=gj.index_params(placement=_lib.PLACE_HOST))
    dp = gj.DeviceArray.from_numpy(probes)
    out = gj.DeviceArray(m, np.uint32)
    for approach, kw in (("partitioning", {"radix_bits": 12}), ("plain", {})):
        if approach == "plain" and kind != "learned_plr":
            continue  # a plain B-tree or Harmonia walk over host memory takes seconds here
        out.fill(0x5A)
        idx.lookup(dp, out, approach=approach, **kw)
        gj.synchronize()
        assert np.array_equal(out.numpy(), exp), approach
    assert idx.stats()["dropped_tuples"] == 0
    idx.close()
    del view
    hcol.free()
