"""The tpch_14 CLI: dbgen .tbl loading and TPC-H query 14 end to end.

The fixture writes small lineitem.tbl / part.tbl files in dbgen's format
(pipe-separated, a trailing '|', two-decimal numerics, ISO dates).  The expected
answer is computed here in plain Python: the date window from the calendar
(Julian day number = proleptic ordinal + 1721425), the sums in exact integers,
the final value with process_results' formula (src/tpch/tpch_14_common.cu:124-127).
The reference's own known answers need dbgen data at SF 1 / SF 10 (not shipped):
the CLI checks them when the row count matches.
"""
import datetime
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpujoin_amd", "tpch_14")

TYPES = ["PROMO BURNISHED COPPER", "STANDARD POLISHED TIN", "PROMO ANODIZED STEEL", "ECONOMY BRUSHED NICKEL",
         "LARGE PLATED BRASS", "PROMO PLATED BRASS", "SMALL ANODIZED COPPER"]


def _ensure():
    if not os.path.exists(CLI):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gpujoin_amd"), "tpch_14"], check=True)


def _jdn(d: datetime.date) -> int:
    return d.toordinal() + 1721425


def _money(cents: int) -> str:
    return f"{cents // 100}.{cents % 100:02d}"


def _write_tables(path, n_part, n_line, seed, shuffle_part=False, crlf=False):
    rng = np.random.default_rng(seed)
    ptype = [TYPES[i] for i in rng.integers(0, len(TYPES), n_part)]
    order = np.arange(n_part)
    if shuffle_part:
        rng.shuffle(order)
    eol = "\r\n" if crlf else "\n"
    with open(os.path.join(path, "part.tbl"), "w", newline="") as f:
        for i in order:
            pk = int(i) + 1
            f.write(f"{pk}|goldenrod lavender spring|Manufacturer#1|Brand#13|{ptype[i]}|7|JUMBO PKG|"
                    f"{_money(90000 + pk % 20000)}|ly. slyly ironi|{eol}")
    first = datetime.date(1995, 7, 15)
    pks = rng.integers(1, n_part + 40, n_line)          # some keys have no part row
    days = rng.integers(0, 120, n_line)
    price = rng.integers(90000, 10500000, n_line)       # hundredths
    disc = rng.integers(0, 11, n_line)                   # hundredths
    with open(os.path.join(path, "lineitem.tbl"), "w", newline="") as f:
        for j in range(n_line):
            ship = first + datetime.timedelta(days=int(days[j]))
            f.write(f"{j // 4 + 1}|{pks[j]}|{(pks[j] * 7) % 1000 + 1}|{j % 4 + 1}|{_money(100 * (j % 50 + 1))}|"
                    f"{_money(int(price[j]))}|0.{int(disc[j]):02d}|0.04|N|O|{ship.isoformat()}|"
                    f"{(ship + datetime.timedelta(days=10)).isoformat()}|"
                    f"{(ship + datetime.timedelta(days=20)).isoformat()}|DELIVER IN PERSON|TRUCK|"
                    f"egular courts above the|{eol}")
    lo, hi = _jdn(datetime.date(1995, 9, 1)), _jdn(datetime.date(1995, 10, 1))
    num = den = hits = 0
    for j in range(n_line):
        d = _jdn(first + datetime.timedelta(days=int(days[j])))
        if not (lo <= d < hi) or pks[j] > n_part:
            continue
        s = int(price[j]) * (100 - int(disc[j]))
        den += s
        hits += 1
        if ptype[pks[j] - 1].startswith("PROMO"):
            num += s
    result = 100 * (num * 1000) // (den // 1000) if den // 1000 else 0
    return num, den, hits, result


def _fixed6(r):
    return f"{r // 1000000}.{r % 1000000:06d}"


def _run(args, timeout=120):
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)


def _host_line(out):
    m = re.search(r"host numerator: (-?\d+) denominator: (-?\d+) matches: (\d+)", out)
    return tuple(int(v) for v in m.groups())


def _wrap64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >> 63 else v


def _c_div(a, b):  # C integer division truncates toward zero
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def test_promo_revenue_closed_form(gj, orc):
    # process_results (src/tpch/tpch_14_common.cu:124-127): 100 * (num*1000) / (den/1000)
    # in int64, whose products wrap for large sums (the last two cases)
    cases = [(0, 0), (1, 999), (16380778, 100000000), (123456789012, 987654321098),
             (70_000_000_000_000_000, 90_000_000_000_000_000), (-(1 << 62), 5_000_000)]
    for num, den in cases:
        d = _c_div(den, 1000)
        want = 0 if d == 0 else _c_div(_wrap64(100 * _wrap64(num * 1000)), d)
        assert gj.promo_revenue(num, den) == want == orc.promo_revenue(num, den), (num, den)
    assert gj.promo_revenue(1, 2) == 0


def test_julian_day_window():
    # the reference's constants (src/tpch/config.hpp:36-37) from the calendar
    assert _jdn(datetime.date(1995, 9, 1)) == 2449962
    assert _jdn(datetime.date(1995, 10, 1)) == 2449992


@pytest.mark.parametrize("shuffle_part,crlf", [(False, False), (True, True)])
def test_host_answer_matches_python(tmp_path, shuffle_part, crlf):
    _ensure()
    num, den, hits, result = _write_tables(str(tmp_path), 3000, 20000, 0x14, shuffle_part, crlf)
    assert hits > 100 and 0 < num < den
    r = _run(["--cpu-only", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    assert "lineitem_size: 20000 part_size: 3000" in r.stdout
    assert _host_line(r.stdout) == (num, den, hits)
    assert f"host query result: {_fixed6(result)}" in r.stdout


def test_host_answer_multithreaded_split(tmp_path):
    # over 1 MiB: the loader splits the file across threads at line boundaries
    _ensure()
    num, den, hits, result = _write_tables(str(tmp_path), 5000, 60000, 0x15)
    assert os.path.getsize(tmp_path / "lineitem.tbl") > (1 << 20)
    r = _run(["--cpu-only", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    assert _host_line(r.stdout) == (num, den, hits)
    assert f"host query result: {_fixed6(result)}" in r.stdout


def test_malformed_table_is_an_error(tmp_path):
    _ensure()
    _write_tables(str(tmp_path), 100, 500, 0x16)
    with open(tmp_path / "lineitem.tbl", "a") as f:
        f.write("1|2|3|4|1.00|2.00|0.05|0.01|N|O|1995-13-45|1995-09-01|1995-09-01|X|Y|z|\n")
    r = _run(["--cpu-only", str(tmp_path)])
    assert r.returncode == 1 and "malformed" in r.stderr


def test_bad_flags(tmp_path):
    _ensure()
    r = _run(["-a", "nested_loops", str(tmp_path)])
    assert r.returncode == 1 and "unknown approach" in r.stderr
    r = _run(["--help"])
    assert r.returncode == 0 and "db_path" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("approach,index", [("hj", "btree"), ("ij_plain", "btree"), ("ij_plain", "learned_plr"),
                                            ("ij_np_lf_lr_plain", "radix_spline"), ("ij_partitioning", "harmonia"),
                                            ("ij_partitioning", "learned_plr"), ("ij_p_bws", "btree"),
                                            ("ij_p_bws", "learned_plr"), ("ij_np_lf_plain", "btree"),
                                            ("ij_np_lf_plain", "harmonia")])
def test_gpu_query14(gpu, tmp_path, approach, index):
    _ensure()
    num, den, hits, result = _write_tables(str(tmp_path), 4096, 30000, 0x17)
    out = tmp_path / "q14.yml"
    print(f"[case] tpch_14 -a {approach} -i {index}", flush=True)
    r = _run(["-a", approach, "-i", index, "-d", str(out), str(tmp_path)])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(f"query result: {_fixed6(result)}") >= 3 + 1  # host line + every repetition
    assert f"numerator: {num} denominator: {den}" in r.stdout
    text = out.read_text()
    assert text.count("name: tpch_query14") == 3
    assert f"approach: {approach}" in text and "rows_per_s:" in text
