"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def _ensure_built(path, make_dir):
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-j8", "-C", make_dir], check=True)


@pytest.fixture(scope="session")
def gj():
    _ensure_built(os.path.join(ROOT, "gpujoin_amd", "libgj.so"), os.path.join(ROOT, "gpujoin_amd"))
    import gpujoin_amd
    gpujoin_amd._lib.load()
    return gpujoin_amd


@pytest.fixture(scope="session")
def orc():
    _ensure_built(os.path.join(ROOT, "oracle", "liborc.so"), os.path.join(ROOT, "oracle"))
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu(gj):
    """The GPU path: fails loudly (no CPU fallback) when no device is visible."""
    n = gj.device_count()
    assert n >= 1, "no GPU visible to libgj"
    info = gj.device_info()
    print(f"[gpu] {info}", flush=True)
    return info
