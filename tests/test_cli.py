"""The index_lookup CLI (reference flags, YAML records)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpujoin_amd", "index_lookup")


def _ensure():
    if not os.path.exists(CLI):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gpujoin_amd"), "index_lookup"], check=True)


def test_cli_help():
    _ensure()
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "-a approach" in r.stdout


def test_cli_rejects_bad_flags():
    _ensure()
    r = subprocess.run([CLI, "-d", "bogus"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "dataset" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("approach,index", [("plain", "btree"), ("partitioning", "harmonia"), ("bws", "radix_spline"),
                                            ("plain", "learned_plr"), ("hj", "binary_search")])
def test_cli_records(gpu, approach, index, tmp_path):
    _ensure()
    out = tmp_path / "r.yml"
    r = subprocess.run([CLI, "-a", approach, "-i", index, "-e", "100000", "-l", "1000000", "-p", "zipf", "-z", "1.0",
                        "-w", "65536", "-r", "3", "-o", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    text = out.read_text()
    assert text.count("name: plain_lookup") == 3  # every repetition validated and recorded
    for field in ("approach: " + approach, "index_type: " + index, "num_lookups: !int64 1000000",
                  "lookups_per_s:", "dt_prefix_sum:", "index_size: !int64", "gather_roofline_frac:",
                  "algorithmic_GBps:"):
        assert field in text
