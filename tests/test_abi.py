"""CPU tests of the C ABI: the library loads, exports every symbol the headers
declare, and rejects bad arguments before touching the device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(gj_[a-z0-9_]+)\s*\(", text))


def test_every_declared_symbol_is_exported(gj):
    lib = gj._lib.load()
    names = _declared("gj.h") | _declared("gj_runtime.h")
    assert len(names) > 40
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, missing
    # and the Python binding knows every one of them
    assert not sorted(names - set(gj._lib.SIGNATURES)), sorted(names - set(gj._lib.SIGNATURES))


def test_parse_index_type_names(gj):
    lib = gj._lib.load()
    for name, kind in gj._lib.INDEX_NAMES.items():
        k = C.c_int(-1)
        assert lib.gj_parse_index_type(name.encode(), C.byref(k)) == 0 and k.value == kind
        assert lib.gj_index_type_name(kind).decode() == name
    k = C.c_int()
    assert lib.gj_parse_index_type(b"bogus", C.byref(k)) == gj._lib.GJ_ERR_INVALID
    assert b"unknown index type" in lib.gj_last_error()


def test_param_defaults(gj):
    p = gj.index_params()
    assert (p.rs_radix_bits, p.rs_max_error, p.harmonia_group_lanes, p.learned_leaf_bits) == (18, 32, 8, 0)
    lp = gj.lookup_params()
    assert lp.approach == 0 and lp.block_size == 256 and lp.ignore_bits == 0xFFFFFFFF


def test_build_rejects_bad_arguments_before_device_use(gj):
    lib = gj._lib.load()
    out = C.c_void_p()
    dummy = C.c_void_p(0x1000)
    # null out, bad key width, empty column, row ids beyond uint32
    assert lib.gj_index_build(1, None, dummy, 10, 8, None, None, None) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_index_build(1, None, dummy, 10, 3, None, None, C.byref(out)) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_index_build(1, None, dummy, 0, 8, None, None, C.byref(out)) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_index_build(1, None, dummy, 1 << 32, 8, None, None, C.byref(out)) == gj._lib.GJ_ERR_OVERFLOW
    assert b"value_t" in lib.gj_last_error()
    assert lib.gj_index_build(1, None, C.c_void_p(0x1004), 10, 8, None, None, C.byref(out)) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_index_build(99, None, dummy, 10, 8, None, None, C.byref(out)) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_index_lookup(None, dummy, 10, dummy, None, None) == gj._lib.GJ_ERR_INVALID


def test_partition_capacity_and_promo_revenue(gj):
    lib = gj._lib.load()
    assert lib.gj_partition_capacity(1000, 8, 4) == 1000 + 16 * 16
    assert lib.gj_partition_capacity(1000, 4, 4) == 1000 + 16 * 32
    assert lib.gj_partition_capacity(1000, 8, 16) == 1000 + 65536 * 16
    assert lib.gj_partition_capacity(1000, 8, 17) == 0
    assert gj.promo_revenue(16380778, 100000000) == 100 * 16380778 * 1000 // (100000000 // 1000)


def test_mlp_entry_points_reject_nulls(gj):
    lib = gj._lib.load()
    assert lib.gj_mlp_loss_grad(None, 32, None, None, 10, None, None, None) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_gather(None, 16, None, 16, None, 256, 4, None) == gj._lib.GJ_ERR_INVALID
    assert lib.gj_gather(C.c_void_p(16), 15, C.c_void_p(16), 16, C.c_void_p(16), 256, 4, None) == \
        gj._lib.GJ_ERR_INVALID


def test_version_string(gj):
    assert b"gfx950" in gj._lib.load().gj_version()
