"""Pins the oracle (oracle/fc5_oracle.c) against reference-held vectors and
against independent mathematics.  CPU only.

Reference vectors used (tests/golden/, written by make_golden.py from the
reference's own sources compiled in oracle/_ref):
  xor_tables.json  xor_build_save output bytes (SHA-256) for several (N, r)
  units_100001_102000.txt   fourcore_classes_v3.hpp units for B in [100001,102000]
  residues.json    free-term residue specs at sampled units up to B=3.52M
  seeds.json       seed counts per class
plus the reference's published Stage-1 count (README.md:146) and self-test
KATs (fourcore_hunt_v3.cpp:49-67, fourcore_cls5_gpu_v4.cu:191-201,229).
"""
import hashlib
import json
import os
import random
import subprocess

import numpy as np
import pytest

import bigint_ref as br
import oracle as O
from conftest import GOLDEN, REF_DIR

TABLES = json.load(open(os.path.join(GOLDEN, "xor_tables.json")))


@pytest.mark.parametrize("rec", TABLES, ids=lambda r: f"N{r['N']}_r{r['r']}")
def test_xor_builder_bytes_match_reference(rec, tmp_path):
    t = O.Table.build_pairs(rec["N"], rec["r"])
    p = str(tmp_path / "t.bin")
    t.save(p)
    b = open(p, "rb").read()
    assert len(b) == rec["bytes"]
    assert b[:64].hex() == rec["header_hex"]
    assert hashlib.sha256(b).hexdigest() == rec["sha256"]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "xor_build_save")), reason="oracle/_ref not built")
def test_xor_builder_matches_compiled_reference_live(tmp_path):
    p = str(tmp_path / "ref.bin")
    subprocess.run([os.path.join(REF_DIR, "xor_build_save"), "--N", "777", "--r", "40", "--out", p],
                   check=True, capture_output=True)
    q = str(tmp_path / "or.bin")
    O.Table.build_pairs(777, 40).save(q)
    assert open(p, "rb").read() == open(q, "rb").read()


def test_xor_no_false_negatives_and_python_query():
    N = 300
    t = O.Table.build_pairs(N, 48)
    hdr, packed = t.header(), t.packed()
    pw = [(x ** 6) % 2 ** 64 for x in range(N + 1)]
    keys = [(pw[i] + pw[j]) % 2 ** 64 for i in range(1, N + 1) for j in range(i, N + 1)]
    assert all(t.contains(k) for k in keys)
    rnd = random.Random(1)
    sample = rnd.sample(keys, 300) + [rnd.getrandbits(64) for _ in range(2000)]
    for k in sample:
        assert t.contains(k) == br.xor_query(hdr, packed, k)


def test_fpr_smoke_threshold():
    assert O.Table.build_pairs(300, 48).fpr_smoke() <= 16      # accepted (fourcore_xor_store.hpp:54-65)
    assert O.Table.build_pairs(300, 8).fpr_smoke() > 16        # refused by the searcher


def test_units_100001_102000_match_reference():
    cls, Bs, us = O.emit_units(100001, 102000, classes=(2, 3, 4, 5))
    mine = "".join(f"{c} {b} {u}\n" for c, b, u in zip(cls, Bs, us))
    assert mine == open(os.path.join(GOLDEN, "units_100001_102000.txt")).read()


def test_residues_match_reference():
    for rec in json.load(open(os.path.join(GOLDEN, "residues.json"))):
        if rec["cls"] != 5:
            continue
        (er, em, ef), (dr, dm, df) = O.free_de_cls5(rec["B"], rec["u"])
        assert (er, em, ef) == (rec["e_res"], rec["e_mod"], rec["e_fmax"])
        assert (dr, dm, df) == (rec["d_res"], rec["d_mod"], rec["d_fmax"])


def test_seed_counts():
    assert [len(O.seeds(100003, c)) for c in (1, 2, 3, 4, 5)] == [144, 24, 36, 6, 6]
    for rec in json.load(open(os.path.join(GOLDEN, "seeds.json"))):
        assert [len(O.seeds(rec["B"], c)) for c in (1, 2, 3, 4, 5)] == rec["counts"]


def test_published_unit_count():
    # README.md:146: units=5765343 for cls5 on B in [4.5M, 5.0M]
    cls, _, _ = O.emit_units(4500000, 5000000, classes=(5,))
    assert len(cls) == 5765343


def test_residue_sets_by_brute_force():
    for B, u in [(100003, 17646), (1000003, 600000 - 600000 % 6), (3200003, 2000004)]:
        (er, _, _), (dr, _, _) = O.free_de_cls5(B, u)
        te = pow(B, 6, 729) * pow(pow(14, 6, 729), -1, 729) % 729
        td = (pow(B, 6, 64) - pow(u, 6, 64)) * pow(pow(21, 6, 64), -1, 64) % 64
        assert er == br.residues(te, 729)
        assert dr == br.residues(td, 64)


def test_pow6_and_mod192_closed_form():
    rnd = random.Random(2)
    xs = [1, 42, 100003, 2353975, 2752470, 2 ** 32 - 1] + [rnd.randrange(1, 2 ** 32) for _ in range(500)]
    for x in xs:
        assert O.pow6_192(x) == x ** 6
    m = 504 * 42 ** 6
    c = 2 ** 64 % m
    for _ in range(500):
        v = rnd.getrandbits(192)
        assert O.mod192(v, m, c) == v % m
        assert O.mod192(v, 247, 2 ** 64 % 247) == v % 247


def test_gate_set_sizes():
    assert len(br.S504) == 27 and len(br.S247) == 50  # fourcore_find_v4.cu:78-88


def test_funnel_identity():
    inv = O.inv216()
    assert (inv * 21 ** 6) % 2 ** 64 == 1
    rnd = random.Random(3)
    for _ in range(2000):
        T = rnd.getrandbits(rnd.choice([60, 100, 125]))
        R = T * 42 ** 6
        fp = ((((R & br.MASK64) >> 6) | (((R >> 64) & br.MASK64) << 58)) * inv) & br.MASK64
        assert fp == T % 2 ** 64


def test_exact_T_kat():
    B, u, d, e = 100003, 17646, 9, 44  # fourcore_cls5_gpu_v4.cu:229
    R = B ** 6 - u ** 6 - (21 * d) ** 6 - (14 * e) ** 6
    assert R % 42 ** 6 == 0
    assert O.T_peel2(B, u, d, e) == R // 42 ** 6
    assert O.T_peel2(B, u, d + 1, e) is None  # not divisible


def _oracle_counts(B, u, table, fmax1=None, fmax2=None, use_gate=True):
    arr, ok = O.pack_units([B], [u])
    assert ok[0]
    if fmax1 is not None:
        arr[0].fmax1 = fmax1
        arr[0].fmax2 = fmax2
    c, hits, _ = O.peel(arr, 1, table, use_gate=use_gate, threads=1)
    return c, [(int(h["e"]), int(h["d"]), int(h["fp"])) for h in hits]


@pytest.mark.parametrize("B,u,fm", [(100003, 17646, None), (101009, None, None),
                                    (3200003, None, (20000, 15000)), (2353975, 85710, (30000, 12000))])
def test_peel_counters_brute_force(B, u, fm):
    """whole-unit counters and maybe-hit list against exact Python residuals"""
    if u is None:  # a real Stage-1 unit at the first B >= B that has one
        _, Bs, us = O.emit_units(B, B + 200, classes=(5,))
        B, u = int(Bs[len(Bs) // 2]), int(us[len(us) // 2])
    table = O.Table.build_pairs(300, 16)  # r=16: a non-empty false-positive list
    f1, f2 = fm if fm else (None, None)
    c, hits = _oracle_counts(B, u, table, f1, f2)
    es, ds = br.cls5_grid(B, u, f1, f2)
    ref, rhits = br.peel_counts(B, u, es, ds, contains=table.contains)
    for k in ref:
        assert c[k] == ref[k], k
    assert sorted(hits) == sorted(rhits)
    c2, _ = _oracle_counts(B, u, table, f1, f2, use_gate=False)
    assert c2["gated"] == 0 and c2["probes"] == c["probes"] + c["gated"]


def test_calls_closed_form_100001_102000():
    _, Bs, us = O.emit_units(100001, 102000, classes=(5,))
    arr, ok = O.pack_units(Bs, us)
    assert ok.all()
    c, _, _ = O.peel(arr, len(Bs), None, threads=8)
    closed = 0
    for b, u in zip(Bs, us):
        es, ds = br.cls5_grid(int(b), int(u))
        closed += len(es) * len(ds)
    assert c["calls"] == closed == 10973276
    assert c["calls"] == c["neg"] + c["zero"] + c["gated"] + c["probes"]


def test_verify_rejects_non_pair_sums():
    B, u, d, e = 100003, 17646, 9, 44
    T = O.T_peel2(B, u, d, e)
    # the true fingerprint of a residual that is not a pair sum: no solution
    assert O.verify_hit(B, u, (B - 1) // 42, d, e, T % 2 ** 64, 2428) == []
    # a fingerprint that is a pair sum mod 2^64 but not equal to T: rejected by the exact check
    assert O.verify_hit(B, u, (B - 1) // 42, d, e, (5 ** 6 + 7 ** 6) % 2 ** 64, 2428) == []


def test_cls234_residues_match_reference():
    for rec in json.load(open(os.path.join(GOLDEN, "residues.json"))):
        if rec["cls"] == 5:
            continue
        res, mod, fmax = O.free_d_cls234(rec["B"], rec["u"], rec["cls"])
        assert (res, mod, fmax) == (rec["d_res"], rec["d_mod"], rec["d_fmax"])
        assert res == br.cls234_residues(rec["cls"], rec["B"], rec["u"])


@pytest.mark.parametrize("cls", [2, 3, 4])
def test_find3_counters_brute_force(cls):
    """jobs, failT, windows, per-gate counters and maybe-hits of whole units"""
    c_, Bs, us = O.emit_units(100001, 100400, classes=(cls,))
    table = O.Table.build_pairs(300, 16)
    for k in (0, len(Bs) // 2):
        B, u = int(Bs[k]), int(us[k])
        oc, oh, _ = O.find3([cls], [B], [u], table, threads=1)
        ref, rhits = br.find3_counts(cls, B, u, contains=table.contains)
        for key in ref:
            assert oc[key] == ref[key], (cls, B, u, key)
        assert sorted((int(h["d"]), int(h["e"]), int(h["fp"])) for h in oh) == sorted(rhits)


def test_find3_exact_T_and_three_core_plant():
    # the peel identity B^6 = u^6 + (f d)^6 + 42^6 T on a real unit
    _, Bs, us = O.emit_units(100001, 100100, classes=(4,))
    B, u = int(Bs[0]), int(us[0])
    res, mod, fmax = O.free_d_cls234(B, u, 4)
    d = res[0]
    T = O.T_peel1(B, u, 7, d)
    assert B ** 6 == u ** 6 + (7 * d) ** 6 + 42 ** 6 * T
    # a fabricated three-core residual T = 3^6 + 4^6 + 5^6 has the window
    # (fourcore_find_v4.cu self-test shape) containing c3 = 5 and its pair 3,4
    T = 3 ** 6 + 4 ** 6 + 5 ** 6
    hi = br.iroot6(T)
    lo = 1
    while lo <= hi and 3 * lo ** 6 < T:
        lo += 1
    assert lo <= 5 <= hi and (T - 5 ** 6) == 3 ** 6 + 4 ** 6
