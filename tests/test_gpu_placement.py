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
