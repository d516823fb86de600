"""The strict residue gate (option "strict_gate"; the reference has no such gate).

Before each xor query the strict gate also requires T mod 64, T mod 81 and
T mod 31, 37, 43, 49, 61, 67, 73, 79 to be sums of two sixth powers.  Pinned here on the CPU:
  - the kernels' residue masks and their row/lane split of T (fc5_strict.hpp,
    compiled for the host) against brute-force pair-sum sets in Python;
  - every sum of two sixth powers passes (no solution can be lost);
  - the oracle's strict peel and find3 against the big-integer brute force of
    whole units: the same counters, `queries` = probes passing the strict gate,
    and the maybe-hit list equal to the unfiltered one with the strict gate
    applied to each hit's exact T.
"""
import os
import random
import subprocess

import pytest

import bigint_ref as br
import oracle as O
from conftest import PKG, REPO


def _mask_set(words, m):
    v = 0
    for i, w in enumerate(words):
        v |= w << (64 * i)
    return {t for t in range(m) if (v >> t) & 1}


def test_kernel_masks_and_row_lane_split(tmp_path):
    exe = str(tmp_path / "strict_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(PKG, "csrc"),
                    os.path.join(REPO, "tests", "native", "strict_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "strict ok" in out.stdout
    masks = {}
    for ln in out.stdout.splitlines():
        if ln.startswith("mask"):
            name, *words = ln.split()
            masks[int(name[4:])] = [int(w, 16) for w in words]
    for m in (31, 37, 43, 64, 81, 49, 61, 67, 73, 79):
        assert _mask_set(masks[m], m) == br.STRICT[m], m


def test_oracle_predicate_matches_sets():
    for t in range(64):
        assert bool(O._lib.or_strict_pass(t, 0, 0, 0, 0)) == (t in br.STRICT[64])
    for t in range(81):
        assert bool(O._lib.or_strict_pass(0, t, 0, 0, 0)) == (t in br.STRICT[81])
    for t in range(31 * 37 * 43):
        want = t % 31 in br.STRICT[31] and t % 37 in br.STRICT[37] and t % 43 in br.STRICT[43]
        assert bool(O._lib.or_strict_pass(0, 0, t, 0, 0)) == want
    for t in range(49 * 61 * 67):
        want = t % 49 in br.STRICT[49] and t % 61 in br.STRICT[61] and t % 67 in br.STRICT[67]
        assert bool(O._lib.or_strict_pass(0, 0, 0, t, 0)) == want
    for t in range(73 * 79):
        want = t % 73 in br.STRICT[73] and t % 79 in br.STRICT[79]
        assert bool(O._lib.or_strict_pass(0, 0, 0, 0, t)) == want


def test_every_pair_sum_passes():
    rnd = random.Random(615)
    for _ in range(20000):
        a, b = rnd.randrange(10 ** 7), rnd.randrange(10 ** 7)
        T = a ** 6 + b ** 6
        assert br.strict_ok(T) and O.strict_pass_T(T)
    # and the gate is selective: about 0.17% of arbitrary residues pass
    frac = sum(br.strict_ok(rnd.getrandbits(100)) for _ in range(100000)) / 100000
    assert 0.001 < frac < 0.003


def _oracle_peel(B, u, table, f1, f2, strict):
    arr, ok = O.pack_units([B], [u])
    assert ok[0]
    if f1 is not None:
        arr[0].fmax1 = f1
        arr[0].fmax2 = f2
    c, hits, _ = O.peel(arr, 1, table, threads=1, strict=strict)
    return c, [(int(h["e"]), int(h["d"]), int(h["fp"])) for h in hits]


@pytest.mark.parametrize("B,u,fm", [(100003, 17646, None), (101009, None, None),
                                    (3200003, None, (20000, 15000)), (2353975, 85710, (30000, 12000))])
def test_oracle_strict_peel_brute_force(B, u, fm):
    if u is None:
        _, Bs, us = O.emit_units(B, B + 200, classes=(5,))
        B, u = int(Bs[len(Bs) // 2]), int(us[len(us) // 2])
    table = O.Table.build_pairs(300, 12)  # r=12: many false positives, so the filtering shows
    f1, f2 = fm if fm else (None, None)
    c, hits = _oracle_peel(B, u, table, f1, f2, True)
    es, ds = br.cls5_grid(B, u, f1, f2)
    ref, rhits = br.peel_counts(B, u, es, ds, contains=table.contains, strict=True)
    for k in ref:
        assert c[k] == ref[k], k
    assert sorted(hits) == sorted(rhits)
    # against the unfiltered run: same gate counters, hits = the strict-passing subset
    c0, hits0 = _oracle_peel(B, u, table, f1, f2, False)
    for k in ("calls", "neg", "zero", "gated", "probes"):
        assert c0[k] == c[k], k
    assert c0["queries"] == 0 and c["queries"] <= c["probes"]
    keep = [h for h in hits0 if O.strict_pass_T(O.T_peel2(B, u, h[1], h[0]))]
    assert sorted(keep) == sorted(hits)


def test_oracle_strict_band_is_the_filtered_default_band():
    """B in [100001,102000]: the strict run's hits are exactly the default run's
    hits whose exact T passes the strict gate; about 3% of the probes query"""
    _, Bs, us = O.emit_units(100001, 102000)
    arr, ok = O.pack_units(Bs, us)
    table = O.Table.build_pairs(2428, 8)  # r=8: enough false positives that some pass the strict gate
    c0, h0, _ = O.peel(arr, len(Bs), table, threads=8)
    c1, h1, _ = O.peel(arr, len(Bs), table, threads=8, strict=True)
    for k in ("calls", "neg", "zero", "gated", "probes"):
        assert c0[k] == c1[k], k
    assert 0.01 * c1["probes"] < c1["queries"] < 0.06 * c1["probes"]
    keep = [(int(h["unit"]), int(h["e"]), int(h["d"])) for h in h0
            if O.strict_pass_T(O.T_peel2(int(Bs[h["unit"]]), int(us[h["unit"]]), int(h["d"]), int(h["e"])))]
    got = [(int(h["unit"]), int(h["e"]), int(h["d"])) for h in h1]
    assert sorted(keep) == sorted(got) and 0 < len(got) < len(h0)


@pytest.mark.parametrize("cls", [2, 3, 4])
def test_oracle_strict_find3_brute_force(cls):
    _, Bs, us = O.emit_units(100001, 100400, classes=(cls,))
    table = O.Table.build_pairs(300, 12)
    for k in (0, len(Bs) // 2):
        B, u = int(Bs[k]), int(us[k])
        oc, oh, _ = O.find3([cls], [B], [u], table, threads=1, strict=True)
        ref, rhits = br.find3_counts(cls, B, u, contains=table.contains, strict=True)
        for key in ref:
            assert oc[key] == ref[key], (cls, B, u, key)
        assert sorted((int(h["d"]), int(h["e"]), int(h["fp"])) for h in oh) == sorted(rhits)
        o0, _, _ = O.find3([cls], [B], [u], table, threads=1)
        assert o0["probes"] == oc["probes"] and oc["queries"] <= oc["probes"]
