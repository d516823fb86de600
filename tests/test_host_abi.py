"""The product's host side through its C ABI and CLIs (no device compute).

Checks that libfc5.so loads and exports every entry point include/fc5.h
declares, that the host logic behind the ABI (Stage-1 emission, closed-form
calls, the byte-identical CPU table builder) matches the reference-derived
golden vectors and the oracle, and that the CLI front ends keep the
reference's flags, outputs and failure behaviour.
"""
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

import bigint_ref as br
import fc5
import oracle as O
from conftest import BIN, GOLDEN, REPO

TABLES = {(r["N"], r["r"]): r for r in json.load(open(os.path.join(GOLDEN, "xor_tables.json")))}


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(REPO, "include", "fc5.h")).read()
    declared = set(re.findall(r"\b(fc5_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations found"
    import ctypes
    lib = ctypes.CDLL(fc5.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared <= set(fc5.EXPORTS)
    assert lib.fc5_abi_version() == 3


def test_emit_units_matches_reference_golden():
    cls, Bs, us = fc5.emit_units(100001, 102000, classes=(2, 3, 4, 5))
    mine = "".join(f"{c} {b} {u}\n" for c, b, u in zip(cls, Bs, us))
    assert mine == open(os.path.join(GOLDEN, "units_100001_102000.txt")).read()


def test_emit_units_matches_oracle_other_ranges():
    for lo, hi in [(1000001, 1000400), (3200001, 3200200), (2353900, 2354100)]:
        a = fc5.emit_units(lo, hi, classes=(2, 3, 4, 5))
        b = O.emit_units(lo, hi, classes=(2, 3, 4, 5))
        for x, y in zip(a, b):
            assert np.array_equal(x.astype(np.int64), y)


def test_units_calls_closed_form():
    _, Bs, us = fc5.emit_units(100001, 102000)
    assert fc5.units_calls(fc5.units_array(Bs, us)) == 10973276
    _, Bs, us = fc5.emit_units(3200001, 3200100)
    closed = sum(len(es) * len(ds) for es, ds in (br.cls5_grid(int(b), int(u)) for b, u in zip(Bs, us)))
    assert fc5.units_calls(fc5.units_array(Bs, us)) == closed
    # either side of B = 14 * 2^21, where e's free range reaches 2^21 (batched path
    # below, per-point 192-bit path above): 9,049,577,666 points, the oracle's count
    pick = [(29360119, 106578), (29360119, 14717196), (29360123, 11148),
            (29360141, 65340), (29360143, 165162), (29360143, 14658612)]
    closed = sum(len(es) * len(ds) for es, ds in (br.cls5_grid(b, u) for b, u in pick))
    assert fc5.units_calls(fc5.units_array([b for b, _ in pick], [u for _, u in pick])) == closed == 9049577666


@pytest.mark.parametrize("N,r", [(300, 48), (2000, 16), (2428, 48)])
def test_build_pairs_file_matches_reference(N, r, tmp_path):
    p = str(tmp_path / "t.bin")
    fc5.build_pairs_file(N, r, p)
    b = open(p, "rb").read()
    assert hashlib.sha256(b).hexdigest() == TABLES[(N, r)]["sha256"]


def test_plant_points_are_exact():
    _, Bs, us = fc5.emit_units(100001, 102000)
    pts = fc5.plant_points(Bs, us, 32, 0x615 << 8 | 1)
    assert len({(i, e, d) for i, e, d, _ in pts}) == 32
    for i, e, d, key in pts:
        assert O.T_peel2(int(Bs[i]), int(us[i]), d, e) % 2 ** 64 == key


def _run(*args, check=True):
    return subprocess.run([os.path.join(BIN, args[0]), *map(str, args[1:])], capture_output=True, text=True,
                          check=check)


def test_cli_xor_build_save_bytes(tmp_path):
    p = str(tmp_path / "x.bin")
    out = _run("xor_build_save", "--N", 300, "--r", 48, "--out", p)
    assert "[xor_build_save] wrote" in out.stdout
    assert hashlib.sha256(open(p, "rb").read()).hexdigest() == TABLES[(300, 48)]["sha256"]
    assert _run("xor_build_save", "--N", 3, check=False).returncode == 1
    assert _run("xor_build_save", "--N", 300, "--r", 7, check=False).returncode == 1


def test_cli_hunt(tmp_path):
    assert "SELFTEST PASS" in _run("fourcore_hunt_v3", "--selftest").stdout
    p = str(tmp_path / "u.buc")
    _run("fourcore_hunt_v3", "--lo", 100001, "--hi", 102000, "--classes", "2,3,4,5", "--emit-units", p)
    assert open(p).read() == open(os.path.join(GOLDEN, "units_100001_102000.txt")).read()
    d = _run("fourcore_hunt_v3", "--density", "--lo", 100001, "--hi", 102000, "--classes", "5").stdout
    assert "cls5  units=615  expanded_jobs~10973276" in d


def test_cli_cls5_host_paths(tmp_path):
    assert "[selftest-host] PASS" in _run("fourcore_cls5_gpu_v4", "--selftest-host").stdout
    p = str(tmp_path / "x.bin")
    fc5.build_pairs_file(300, 48, p)
    ok = _run("fourcore_cls5_gpu_v4", "--check-table", "--N", 300, "--load-table", p)
    assert "--check-table OK N=300" in ok.stderr
    bad = _run("fourcore_cls5_gpu_v4", "--check-table", "--N", 301, "--load-table", p, check=False)
    assert bad.returncode == 1 and "N mismatch" in bad.stderr and "FATAL" in bad.stderr
    q = str(tmp_path / "r8.bin")
    fc5.build_pairs_file(300, 8, q)
    fpr = _run("fourcore_cls5_gpu_v4", "--check-table", "--N", 300, "--load-table", q, check=False)
    assert fpr.returncode == 1 and "FPR smoke too high" in fpr.stderr
    assert _run("fourcore_cls5_gpu_v4", "--units", "/nonexistent.buc", check=False).returncode == 1
    # a damaged header (cell count far beyond the file) is refused before anything is allocated
    raw = bytearray(open(p, "rb").read())
    raw[32:40] = (0x1FFFFFFFFFFFFFFF).to_bytes(8, "little")  # m_cells
    d = str(tmp_path / "damaged.bin")
    open(d, "wb").write(bytes(raw))
    dmg = _run("fourcore_cls5_gpu_v4", "--check-table", "--N", 300, "--load-table", d, check=False)
    assert dmg.returncode == 1 and "short read" in dmg.stderr


def test_cli_617_argument_refusals():
    """solve_617_v1 runs Branch A class 5 only; everything else is refused
    before any device use (exit 1, as the reference does for bad arguments)"""
    for args in [("1000", "2000", "1,5", "--branch-a-only"), ("1000", "2000", "5"),
                 ("1000", "2400000", "5", "--branch-a-only"), ("1000", "2000", "5", "--branch-a-only", "--xcheck"),
                 ("1000", "2000", "5", "--branch-a-only", "--bogus"), ("100", "200", "5", "--branch-a-only")]:
        r = _run("solve_617_v1", *args, check=False)
        assert r.returncode == 1, args
    assert "usage:" in _run("solve_617_v1", check=False).stderr


def _bbj(path):
    import struct
    b = open(path, "rb").read()
    magic, version, n, rec, lo, hi, pad = struct.unpack("<IIQIIII", b[:32])
    assert (magic, version, rec) == (0x4A424346, 2, 36) and len(b) == 32 + 36 * n
    recs = [struct.unpack("<B3xIIIIQQ", b[32 + 36 * i:68 + 36 * i]) for i in range(n)]
    return (n, lo, hi), recs


def test_cli_hunt_bbj_units_and_jobs(tmp_path):
    """.bbj outputs (fourcore_job_bin.hpp layout) carry the same units and jobs
    as the text outputs"""
    p = str(tmp_path / "u.bbj")
    _run("fourcore_hunt_v3", "--lo", 100001, "--hi", 102000, "--classes", "2,3,4,5", "--emit-units", p)
    (n, lo, hi), recs = _bbj(p)
    lines = [f"{c} {B} {u}" for c, B, u, f1, f2, tl, th in recs]
    assert lines == open(os.path.join(GOLDEN, "units_100001_102000.txt")).read().split("\n")[:-1]
    assert all(f1 == f2 == tl == th == 0 for _, _, _, f1, f2, tl, th in recs)
    assert (lo, hi) == (min(r[1] for r in recs), max(r[1] for r in recs))
    tb, tt = str(tmp_path / "j.bbj"), str(tmp_path / "j.but")
    _run("fourcore_hunt_v3", "--lo", 100001, "--hi", 100300, "--classes", "4,5", "--expand", "--emit", tb)
    _run("fourcore_hunt_v3", "--lo", 100001, "--hi", 100300, "--classes", "4,5", "--expand", "--emit", tt)
    _, jobs = _bbj(tb)
    text = [tuple(map(int, l.split())) for l in open(tt).read().split("\n") if l]
    assert [(c, B, u, f1, f2, tl, th) for c, B, u, f1, f2, tl, th in jobs] == text
    assert len(text) > 1000


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits non-zero with the reason"""
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(repo, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


@pytest.mark.parametrize("r", [12, 24, 40, 64])
def test_host_builder_other_widths_match_oracle_and_reference(r, tmp_path):
    """cells of 12, 24 and 40 bits straddle byte and word boundaries, 64 bits
    fill the word: the product's builder writes the oracle's bytes, and the
    reference xor_build_save's where oracle/_ref is built"""
    import subprocess
    p = str(tmp_path / "mine.bin")
    fc5.build_pairs_file(500, r, p)
    q = str(tmp_path / "oracle.bin")
    O.Table.build_pairs(500, r).save(q)
    mine = open(p, "rb").read()
    assert mine == open(q, "rb").read()
    ref = os.path.join(REPO, "oracle", "_ref", "xor_build_save")
    if os.path.exists(ref):
        s = str(tmp_path / "ref.bin")
        subprocess.run([ref, "--N", "500", "--r", str(r), "--out", s], check=True, capture_output=True)
        assert mine == open(s, "rb").read()


def test_gate_residue_reduction_in_device_header():
    """the device computes c^6 mod 124488 in 32-bit steps (mulmod_gate in
    fc5_device.hpp): its constant must be 2^32 mod 124488, and the reduction,
    emulated with 32-bit wrap-around, equals direct modular powers at the edges
    of its domain (c < 2^25) and on a random sample"""
    import random
    src = open(os.path.join(REPO, "six-one-five_amd", "csrc", "fc5_device.hpp")).read()
    m = re.search(r"\(u32\)\(p >> 32\) \* (\d+)u", src)
    assert m and int(m.group(1)) == 2 ** 32 % 124488
    k = int(m.group(1))

    def mm(x, y):
        p = x * y
        assert p < 2 ** 51
        r = ((p & 0xFFFFFFFF) % 124488 + (p >> 32) * k) & 0xFFFFFFFF
        return r % 124488

    rnd = random.Random(6)
    cs = list(range(0, 2048)) + [2 ** 25 - 1 - i for i in range(2048)] + [rnd.randrange(2 ** 25) for _ in range(20000)]
    for c in cs:
        c2 = mm(c, c)
        assert mm(mm(c2, c), mm(c2, c)) == pow(c, 6, 124488), c


def test_bench_default_config_finds_its_counter_summaries():
    """the default bench line (config 4) fills roofline.traffic and roofline.valu
    from the committed rocprofv3 summaries named for its band"""
    import bench
    cfg = bench.CONFIGS[4]
    band = f"B{cfg['lo']}-{cfg['hi']}"
    t = bench.load_profile(f"traffic_cls5_batched_{band}.json")
    v = bench.load_profile(f"valu_cls5_batched_{band}.json")
    assert t and t["traffic_bytes_per_launch"] > t["algorithmic_bytes_per_launch"] > 0
    assert 1.9 < t["calibration"]["fetch_factor"] < 2.1
    assert v and 0 < v["VALUBusy"] < 100 and v["SQ_INSTS_VALU"] > v["SQ_INSTS_VALU_INT64"] > 0
    roof = bench.valu_roofline(v, {"valu_wave_instr_per_s": {"mix_7pct_int64": 6e11}})
    assert 0 < roof["frac"] < 1 and 0 < roof["int64_share"] < 0.2


def test_bench_partition_figures_and_lookup_traffic_summary():
    """the partitioned run's roofline (k_part_lookup: event-timed stage sums over
    the pieces, 35 B per probe over a piece's two lookup launches plus the
    staged slices) and the committed traffic summary it reads"""
    import bench

    class H:
        m_cells = 3585713259

    stage = dict(first=128.0, fine0=137.0, lookup0=274.0, fine1=191.0, lookup1=299.0, fine2=191.0, last=65.0,
                 check=0.1, emit=190.0, pipeline=1287.0, identify=257.0)
    cnt = dict(pieces=12, accepted=18, full_queries=518257, chunks=8334171)
    probes = 34008059307
    f = bench.partition_figures((stage, cnt), {"probes": probes}, H(), 1737.0)
    roof = f["roofline"]
    assert roof["kernel"] == "k_part_lookup" and roof["launches_per_step"] == 24
    assert abs(roof["avg_launch_ms"] - (274.0 + 299.0) / 24) < 1e-9
    slices = -(-(H.m_cells // 3) // 65536)
    assert roof["algorithmic_bytes_per_launch"] == (probes * 35 + 24 * slices * 65536) / 24
    assert 0 < roof["frac"] < 1 and f["stages"]["fine1"]["bytes"] == probes * 18
    cfg = bench.CONFIGS[4]
    t = bench.load_profile(f"traffic_part_lookup_B{cfg['lo']}-{cfg['hi']}.json")
    assert t and 1.0 <= t["traffic_bytes_per_launch"] / t["algorithmic_bytes_per_launch"] < 1.2
    assert set(t["pipeline_kernels"]) == {"k_part_first", "k_part_fine", "k_part_lookup", "k_part_last"}


def test_bench_runs_no_reference_binary():
    """bench.py never executes a program built from the reference's sources: the
    reference CPU figure is read from oracle/ref_cpu_timing.json, written by
    oracle/time_ref_cpu.py where the reference sources are"""
    src = open(os.path.join(REPO, "bench.py")).read()
    assert "subprocess" not in src and "solve_516_v2\"" not in src and "_ref\"" not in src
    ref = json.load(open(os.path.join(REPO, "oracle", "ref_cpu_timing.json")))
    assert set(ref["configs"]) == {"1", "2"} and ref["configs"]["1"]["grid_points"] == 10973276


def test_noalloc_limits_match_header():
    """the Python mirror's limits of fc5_run_units_noalloc are the header's"""
    hdr = open(os.path.join(REPO, "include", "fc5.h")).read()
    val = {k: int(v) for k, v in re.findall(r"#define (FC5_NOALLOC_\w+) (\d+)", hdr)}
    assert (val["FC5_NOALLOC_UNITS"], val["FC5_NOALLOC_HITS"], val["FC5_NOALLOC_BMAX"]) == \
        (fc5.NOALLOC_UNITS, fc5.NOALLOC_HITS, fc5.NOALLOC_BMAX)
    assert (fc5.NOALLOC_BMAX - 1) // 21 == 2 ** 18 - 1 and fc5.NOALLOC_BMAX // 21 == 2 ** 18  # B + 1: d past 2^18 - 1
    assert val["FC5_NOALLOC_BYTES"] <= 8 << 20
