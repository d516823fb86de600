"""(6,1,7) class-5 twin (solve_617_v1.cu Branch A, class 5): the oracle and the
library's host generator against the reference's own outputs and mathematics.
CPU only.

Pinned by:
  twin_cands.json      candidate lists written by the reference generator
                       (gen_class_cands, solve_617_v1.cu:885-906) compiled in
                       oracle/_ref, and the reference triple-sum sets (:334-346)
  known_617_cls5.json  the 54 class-5 rows of the reference's list of known
                       (6,1,7) solutions (617-solutions-clean.txt): whole-B runs
                       must recover them
  bigint_ref           a big-integer restatement of the k_find4_cls5 loops from
                       their definitions (every c4 once), on small candidates
  planted identities   base = (21d)^6 + 42^6 (c1^6+c2^6+c3^6+c4^6), the reference's
                       GPU plant pattern (solve_617_v1.cu:1564-1603)
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import bigint_ref as br
import fc5
import oracle as O
from conftest import GOLDEN

TWIN = json.load(open(os.path.join(GOLDEN, "twin_cands.json")))
KNOWN = json.load(open(os.path.join(GOLDEN, "known_617_cls5.json")))


def cand_lines(c):
    out = []
    for x in c:
        base = int(x["base_lo"]) | int(x["base_hi"]) << 64
        assert base < 2 ** 127
        head = [x["B"], x["u"], x["w"], int(x["base_hi"]), int(x["base_lo"]), x["lim"], x["fmax1"], x["mod1"],
                x["nres1"]]
        out.append(" ".join(str(int(v)) for v in head + list(x["res"][:x["nres1"]])))
    return out


@pytest.mark.parametrize("rec", TWIN["cands"], ids=lambda r: f"B{r['B']}_{r['u_band'][0]}_{r['u_band'][1]}")
def test_twin_cands_match_reference(rec):
    for gen in (O.twin_cands, fc5.twin_cands):
        lines = cand_lines(gen(rec["B"], tuple(rec["u_band"])))
        assert len(lines) == rec["count"]
        assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == rec["sha256"]
        if "lines" in rec:
            assert lines == rec["lines"]


def test_twin_cand_identities():
    """u^6 + w^6 + base = B^6 and the role congruences (solve_617_v1.cu:1403-1451)"""
    for B in (54347, 100003, 400471, 2353003):
        c = O.twin_cands(B)
        assert len(c)
        for x in c[:: max(1, len(c) // 50)]:
            base = int(x["base_lo"]) | int(x["base_hi"]) << 64
            u, w = int(x["u"]), int(x["w"])
            assert u ** 6 + w ** 6 + base == B ** 6
            assert u % 14 == 0 and u % 3 and w % 6 == 0 and w % 7
            for r in x["res"][:x["nres1"]]:
                assert pow(21 * int(r), 6, 64) == pow(B, 6, 64)


def test_tri_sets_match_reference_and_definition():
    t504, t247, n504, n247 = O.tri_sets()
    ref = TWIN["tri"]
    assert (n504, n247) == (ref[0], ref[1]) == (64, len(br.T247))
    assert [int(v) for v in t504] == ref[2:10] and [int(v) for v in t247] == ref[10:14]
    assert {x for x in range(504) if (int(t504[x >> 6]) >> (x & 63)) & 1} == br.T504
    assert {x for x in range(247) if (int(t247[x >> 6]) >> (x & 63)) & 1} == br.T247


def small_cand(B, u, w, fmax1, lim=None, res=None):
    base = B ** 6 - u ** 6 - w ** 6
    if res is None:
        res = [r for r in range(64) if r % 2 and pow(21 * r, 6, 64) == pow(B, 6, 64)]
    lim = (B - 1) // 42 if lim is None else lim
    return O.twin_cand(B, u, w, base, lim, fmax1, res), dict(base=base, lim=lim, fmax1=fmax1, res=res)


@pytest.mark.parametrize("c4_div,use_gate,use_tri", [(3, True, True), (4, True, True), (3, True, False),
                                                     (3, False, True)])
def test_twin_oracle_matches_bigint_restatement(c4_div, use_gate, use_tri):
    N = 700
    tab = O.Table.build_pairs(N, 16)  # r = 16: false positives make the hit list non-trivial
    rnd = random.Random(617)
    k = KNOWN[0]
    cases = [small_cand(k["B"], k["u"], k["w"], fmax1=k["d"], res=[k["d"] % 64])]
    for B in (54347, 63631):  # other candidates of the same B, short grids
        c = O.twin_cands(B)
        x = c[rnd.randrange(len(c))]
        cases.append(small_cand(B, int(x["u"]), int(x["w"]), fmax1=rnd.randrange(100, 160)))
    for cand, spec in cases:
        got, hits, nh = O.twin_run(cand, tab, use_gate=use_gate, use_tri=use_tri, c4_div=c4_div)
        want, whits = br.twin_counts(spec, contains=tab.contains, c4_div=c4_div, use_gate=use_gate,
                                     use_tri=use_tri)
        assert got == want
        assert [(int(h["d"]), int(h["c4"]), int(h["c3"]), int(h["fp"])) for h in hits] == sorted(whits)


def test_twin_oracle_strict_matches_bigint_restatement():
    """the strict residue gate on the twin's probe keys V = T4 - c3^6 (an option of
    this build): the oracle against the big-integer restatement, and against the
    default run (same counters but queries; hits = the default hits whose exact V
    passes the strict gate; the known solution's key passes)"""
    N = 700
    tab = O.Table.build_pairs(N, 8)  # r = 8: false positives survive the strict gate too
    rnd = random.Random(6171)
    k = KNOWN[0]
    cases = [small_cand(k["B"], k["u"], k["w"], fmax1=k["d"], res=[k["d"] % 64])]
    for B in (54347, 63631):
        c = O.twin_cands(B)
        x = c[rnd.randrange(len(c))]
        cases.append(small_cand(B, int(x["u"]), int(x["w"]), fmax1=rnd.randrange(100, 160)))
    for cand, spec in cases:
        got, hits, _ = O.twin_run(cand, tab, strict=True)
        want, whits = br.twin_counts(spec, contains=tab.contains, strict=True)
        assert got == want
        assert [(int(h["d"]), int(h["c4"]), int(h["c3"]), int(h["fp"])) for h in hits] == sorted(whits)
        base, bhits, _ = O.twin_run(cand, tab)
        for key in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes"):
            assert base[key] == got[key], key
        assert base["queries"] == 0 and 0 < got["queries"] < got["calls"]

        def V(h):
            T = (spec["base"] - (21 * int(h["d"])) ** 6) // br.M42
            return T - int(h["c4"]) ** 6 - int(h["c3"]) ** 6
        keep = [(int(h["d"]), int(h["c4"]), int(h["c3"])) for h in bhits if br.strict_ok(V(h))]
        assert keep == [(int(h["d"]), int(h["c4"]), int(h["c3"])) for h in hits]
    # the known solution: its pair key is found under the strict gate
    kc, _ = cases[0]
    found = O.twin_run(kc, O.Table.build_pairs(N, 48), strict=True)[1]
    assert any((int(h["d"]), int(h["c4"]), int(h["c3"])) == (k["d"], k["cores"][3], k["cores"][2]) for h in found)


@pytest.mark.parametrize("tile", [16, 60, 2048])
def test_twin_oracle_reference_tiles_match_bigint_restatement(tile):
    """the reference launch's unbounded c4 tiles (solve_617_v1.cu:607-608,961-979):
    block z runs from lo4 + z * tile to hi4, so c4 values far into a window are
    counted (and hit) once per block that reaches them; small tiles make the
    duplicates appear at small B, tile 16 also reaches the 32-block cap"""
    N = 700
    tab = O.Table.build_pairs(N, 16)
    k = KNOWN[0]
    rnd = random.Random(618)
    c = O.twin_cands(63631)
    x = c[rnd.randrange(len(c))]
    cases = [small_cand(k["B"], k["u"], k["w"], fmax1=k["d"], res=[k["d"] % 64]),
             small_cand(63631, int(x["u"]), int(x["w"]), fmax1=100)]
    for cand, spec in cases:
        zmax = min(32, max(1, -(-max(spec["lim"], 1) // tile)))
        got, hits, nh = O.twin_run(cand, tab, tile=tile)
        want, whits = br.twin_counts(spec, contains=tab.contains, tile=tile, zmax=zmax)
        assert got == want
        assert [(int(h["d"]), int(h["c4"]), int(h["c3"]), int(h["fp"])) for h in hits] == sorted(whits)
        once, _, _ = O.twin_run(cand, tab)
        if tile == 2048:  # windows below 2048 at these B: the same as visiting each c4 once
            assert got == once
        else:
            assert got["c4_total"] > once["c4_total"] and got["calls"] > once["calls"]


def test_twin_oracle_finds_known_solutions_whole_B():
    """every candidate of B, table N = B/42 (the reference's N = B_max/42)"""
    for k in [k for k in KNOWN if k["B"] < 70000]:
        B = k["B"]
        N = B // 42
        tab = O.Table.build_pairs(N, 48)
        cands = O.twin_cands(B)
        cnt, hits, nh = O.twin_run(cands, tab)
        assert cnt["maybe"] == nh and nh >= 1
        sols = set()
        for h in hits:
            for terms, g in O.twin_verify(cands[h["cand"]], int(h["d"]), int(h["c4"]), int(h["c3"]),
                                          int(h["fp"]), N):
                if g == 1:
                    sols.add(tuple(terms))
        assert sols == {tuple(k["terms"])}, (B, sols)


def planted(N, n, seed):
    """the reference's find4 plant: fabricated candidates whose T is a sum of
    four distinct sixth powers of cores <= N (solve_617_v1.cu:1564-1579)"""
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        d = 1 + rnd.randrange(280)
        while True:
            cs = [1 + rnd.randrange(N) for _ in range(4)]
            if len(set(cs)) == 4:
                break
        base = (21 * d) ** 6 + br.M42 * sum(x ** 6 for x in cs)
        out.append((O.twin_cand(42 * N + 1, 14, 6, base, N, d, [d % 64]), d, sorted(cs)))
    return out


def test_twin_oracle_plants_exact_and_reference_windows():
    N = 400
    tab = O.Table.build_pairs(N, 48)
    plants = planted(N, 48, 617)
    cands = np.concatenate([p[0] for p in plants])
    for c4_div in (3, 4):
        cnt, hits, nh = O.twin_run(cands, tab, c4_div=c4_div)
        found = set()
        for h in hits:
            _, d, cs = plants[h["cand"]]
            T = sum(x ** 6 for x in cs)
            if int(h["d"]) == d and (T - int(h["c4"]) ** 6 - int(h["c3"]) ** 6) % 2 ** 64 == int(h["fp"]):
                found.add(int(h["cand"]))
        # the plant is recovered iff its largest core is in the c4 window
        want = {i for i, (_, d, cs) in enumerate(plants) if c4_div * cs[3] ** 6 >= sum(x ** 6 for x in cs)}
        assert found >= want
        if c4_div == 4:
            assert want == set(range(len(plants)))
