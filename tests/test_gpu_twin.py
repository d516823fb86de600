"""GPU: the (6,1,7) class-5 twin (libfc5.so, k_twin_batched) against the oracle
and against the reference's list of known (6,1,7) solutions.

Bit-exact bar: counters (jobs, c4_total, c4_tri_skipped, calls, gated, probes,
maybe) and the sorted maybe-hit list equal the oracle's on the same table
bytes; every known class-5 solution is recovered, and whole-B runs report no
primitive solution outside the list.
"""
import json
import os
import sys

import numpy as np
import pytest

import fc5
import oracle as O
from conftest import GOLDEN
from test_twin import KNOWN, planted

pytestmark = pytest.mark.gpu

COUNTERS = ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes", "maybe")
THREADS = min(16, os.cpu_count() or 8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    yield c
    c.close()


def device_table(ctx, N, r):
    """device-built table, its bytes shared with the oracle"""
    ctx.table_build(N, r, on_gpu=True)
    h, packed = ctx.table_download()
    return O.Table.from_bytes({f: getattr(h, f) for f, _ in O.Header._fields_}, packed)


def compare(res, hits, oc, ohits):
    for k in COUNTERS:
        assert res[k] == oc[k], (k, res[k], oc[k])
    got = np.sort(hits, order=["cand", "d", "c4", "c3"])
    assert len(got) == len(ohits)
    for f in ("cand", "d", "c4", "c3", "fp"):
        assert np.array_equal(got[f], ohits[f]), f


def set_opts(ctx, c4_div, use_gate, use_tri):
    ctx.set_option("twin_c4_div", c4_div)
    ctx.set_option("use_gate", int(use_gate))
    ctx.set_option("twin_tri_gate", int(use_tri))


@pytest.mark.parametrize("r", [16, 48])
@pytest.mark.parametrize("c4_div,use_gate,use_tri", [(3, True, True), (4, True, True), (3, True, False),
                                                     (3, False, True)])
def test_twin_whole_B_matches_oracle(ctx, r, c4_div, use_gate, use_tri):
    say(f"twin parity r={r} c4_div={c4_div} gate={use_gate} tri={use_tri}")
    t = device_table(ctx, 1500, r)
    cands = np.concatenate([O.twin_cands(54347), O.twin_cands(59819)])
    set_opts(ctx, c4_div, use_gate, use_tri)
    try:
        res, hits, sols = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
    finally:
        set_opts(ctx, 3, True, True)
    oc, oh, _ = O.twin_run(cands, t, use_gate=use_gate, use_tri=use_tri, c4_div=c4_div, threads=THREADS)
    compare(res, hits, oc, oh)
    assert res["cands"] == len(cands) and res["hits_dropped"] == 0
    if r == 16:
        assert oc["maybe"] > 0


@pytest.mark.parametrize("c4_div,use_gate,use_tri", [(3, True, True), (4, True, True), (3, False, False)])
def test_twin_plain_form_matches_oracle(ctx, c4_div, use_gate, use_tri):
    """the reference-shaped kernel (block per candidate x d x c4 tile) over the same
    candidates and table bytes"""
    say(f"twin plain form c4_div={c4_div} gate={use_gate} tri={use_tri}")
    t = device_table(ctx, 1500, 16)
    cands = np.concatenate([O.twin_cands(54347), O.twin_cands(59819)])
    set_opts(ctx, c4_div, use_gate, use_tri)
    ctx.set_option("twin_plain", 1)
    try:
        res, hits, sols = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
    finally:
        ctx.set_option("twin_plain", 0)
        set_opts(ctx, 3, True, True)
    oc, oh, _ = O.twin_run(cands, t, use_gate=use_gate, use_tri=use_tri, c4_div=c4_div, threads=THREADS)
    compare(res, hits, oc, oh)


def test_twin_large_candidates_match_oracle(ctx):
    """two full candidates at B = 400471 (one holds the known solution), r = 48,
    N = 9535: 4.4e8 c3 values and a 400 MB table"""
    say("twin parity B=400471")
    t = device_table(ctx, 9535, 48)
    c = O.twin_cands(400471)
    pick = np.concatenate([c[(c["u"] == 355502) & (c["w"] == 144756)], c[len(c) // 3:len(c) // 3 + 1]])
    res, hits, sols = ctx.twin_run(pick.view(fc5.TWIN_CAND_DTYPE))
    oc, oh, _ = O.twin_run(pick, t, threads=THREADS)
    compare(res, hits, oc, oh)
    assert [s["terms"] for s in sols] == [[13860, 83202, 93240, 129696, 144756, 355502, 357609]]


@pytest.mark.parametrize("c4_div", [3, 4])
def test_twin_recovers_every_known_solution(ctx, c4_div):
    """the candidate (B, u, w) of each of the 54 known class-5 solutions"""
    say(f"twin known answers c4_div={c4_div}")
    N = max(k["B"] for k in KNOWN) // 42
    ctx.table_build(N, 48, on_gpu=True)
    rows = []
    for k in KNOWN:
        c = fc5.twin_cands(k["B"])
        sel = c[(c["u"] == k["u"]) & (c["w"] == k["w"])]
        assert len(sel) == 1, k["B"]
        rows.append(sel)
    ctx.set_option("twin_c4_div", c4_div)
    try:
        res, hits, sols = ctx.twin_run(np.concatenate(rows))
    finally:
        ctx.set_option("twin_c4_div", 3)
    got = sorted(tuple(s["terms"]) for s in sols if s["gcd"] == 1)
    assert got == sorted(tuple(k["terms"]) for k in KNOWN)


def test_twin_strict_gate_matches_oracle_and_keeps_the_known(ctx):
    """option strict_gate on the twin's keys V = T - c4^6 - c3^6: two whole B
    against the oracle's strict mode (r = 8, so that false positives pass the
    strict gate too), the 54 known solutions all recovered with it on, and the
    plain form refusing it"""
    say("twin strict gate")
    t = device_table(ctx, 1500, 8)
    cands = np.concatenate([O.twin_cands(54347), O.twin_cands(59819)])
    ctx.set_option("strict_gate", 1)
    try:
        res, hits, sols = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
        oc, oh, _ = O.twin_run(cands, t, strict=True, threads=THREADS)
        compare(res, hits, oc, oh)
        assert res["strict_out"] == oc["calls"] - oc["queries"] and 0 < oc["queries"] < oc["calls"]
        assert len(oh) > 0
        N = max(k["B"] for k in KNOWN) // 42
        ctx.table_build(N, 48, on_gpu=True)
        rows = []
        for k in KNOWN:
            c = fc5.twin_cands(k["B"])
            rows.append(c[(c["u"] == k["u"]) & (c["w"] == k["w"])])
        res, hits, sols = ctx.twin_run(np.concatenate(rows))
        assert sorted(tuple(s["terms"]) for s in sols if s["gcd"] == 1) == sorted(tuple(k["terms"]) for k in KNOWN)
        assert res["strict_out"] > 0
        ctx.set_option("twin_plain", 1)
        with pytest.raises(fc5.Fc5Error):
            ctx.twin_run(np.concatenate(rows))
    finally:
        ctx.set_option("strict_gate", 0)
        ctx.set_option("twin_plain", 0)


def test_twin_whole_B_ranges_find_exactly_the_known(ctx):
    """every candidate of every Branch-A B of small ranges around known B"""
    say("twin whole-B ranges")
    known_B = {k["B"]: tuple(k["terms"]) for k in KNOWN}
    ctx.table_build(130000 // 42, 48, on_gpu=True)
    for lo, hi in [(54300, 54400), (97000, 97100), (123900, 124000)]:
        res, hits, sols = ctx.twin_run_range(lo, hi)
        got = {tuple(s["terms"]) for s in sols if s["gcd"] == 1}
        want = {t for B, t in known_B.items() if lo <= B <= hi}
        assert got == want, (lo, hi, got)
        per_B = [fc5.twin_cands(B) for B in range(lo, hi + 1) if B % 2 and B % 3 and B % 7]
        assert res["cands"] == sum(len(c) for c in per_B)


def test_twin_plants_recovered(ctx):
    """the reference's find4 plant pattern (solve_617_v1.cu:1564-1603), 256 plants"""
    say("twin plants")
    N = 2000
    ctx.table_build(N, 48, on_gpu=True)
    plants = planted(N, 256, 1617)
    cands = np.concatenate([p[0] for p in plants]).view(fc5.TWIN_CAND_DTYPE)
    for c4_div in (4, 3):
        ctx.set_option("twin_c4_div", c4_div)
        try:
            res, hits, _ = ctx.twin_run(cands)
        finally:
            ctx.set_option("twin_c4_div", 3)
        found = set()
        for h in hits:
            _, d, cs = plants[h["cand"]]
            T = sum(x ** 6 for x in cs)
            if int(h["d"]) == d and (T - int(h["c4"]) ** 6 - int(h["c3"]) ** 6) % 2 ** 64 == int(h["fp"]):
                found.add(int(h["cand"]))
        want = {i for i, (_, d, cs) in enumerate(plants) if c4_div * cs[3] ** 6 >= sum(x ** 6 for x in cs)}
        assert found >= want
        if c4_div == 4:
            assert len(found) == 256


def test_twin_errors(ctx):
    ctx.table_build(500, 8, on_gpu=True)
    c = fc5.twin_cands(54347)
    bad = c[:1].copy()
    bad["base_lo"] += 1  # base - (21d)^6 no longer a multiple of 42^6
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.twin_run(bad)
    assert e.value.code == fc5.E_ARG
    with pytest.raises(fc5.Fc5Error) as e:  # B^6 < 2^127 cap, solve_617_v1.cu:146,1775
        ctx.twin_run_range(2353973, 2353974)
    assert e.value.code == fc5.E_ARG
    ctx.set_option("hit_cap", 4)  # r = 8: many maybe-hits
    try:
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.twin_run(c)
        assert e.value.code == fc5.E_OVERFLOW
        res, hits, _ = ctx.twin_run(c, allow_overflow=True)
        assert res["hits_dropped"] == res["n_hits"] - 4 and len(hits) == 4
    finally:
        ctx.set_option("hit_cap", 1 << 20)


def test_cli_617_finds_and_checks_known(tmp_path):
    """the drop-in CLI over [54000, 64000]: SOLUTION lines of the known class-5
    rows, --check-known against a list in the reference's file format"""
    import subprocess
    from conftest import BIN
    known = tmp_path / "known.txt"
    known.write_text("# known list\n" + "".join(
        f"{k['B']} = {' + '.join(str(x) for x in sorted(k['terms'], reverse=True))}  | 2nd kind | ref\n"
        for k in KNOWN))
    r = subprocess.run([os.path.join(BIN, "solve_617_v1"), "54000", "64000", "5", "--branch-a-only", "--quiet",
                        "--check-known", "--known-file", str(known), "--counters"],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    sols = [l for l in r.stdout.splitlines() if l.startswith("SOLUTION") and "primitive" in l and "NON" not in l]
    want = {k["B"] for k in KNOWN if 54000 <= k["B"] <= 64000}
    assert {int(l.split("B=")[1].split()[0]) for l in sols} == want
    assert "[check-known]" in r.stderr and "PASS" in r.stderr
    assert any(l.startswith("COUNTERS cands=") for l in r.stdout.splitlines())


@pytest.mark.parametrize("tile", [16, 60])
@pytest.mark.parametrize("plain", [0, 1])
def test_twin_reference_tiles_small_B(ctx, tile, plain):
    """the reference launch's unbounded c4 tiles (twin_ref_tiles) at small tile
    widths, so that repeated visits (and the 32-tile cap, at 16) occur at small
    B: counters and the maybe-hit list, duplicates included, equal the oracle's
    restatement of the reference loops, in both kernel forms"""
    say(f"twin reference tiles={tile} plain={plain}")
    t = device_table(ctx, 1500, 16)
    cands = np.concatenate([O.twin_cands(54347)[::3], O.twin_cands(59819)[::5]])
    ctx.set_option("twin_ref_tiles", tile)
    ctx.set_option("twin_plain", plain)
    try:
        res, hits, _ = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
    finally:
        ctx.set_option("twin_ref_tiles", 0)
        ctx.set_option("twin_plain", 0)
    oc, oh, _ = O.twin_run(cands, t, threads=THREADS, tile=tile)
    compare(res, hits, oc, oh)
    once, _, _ = O.twin_run(cands, t, threads=THREADS)
    assert oc["calls"] > once["calls"] and oc["maybe"] > once["maybe"] > 0


def test_twin_reference_tiles_2048_past_515k(ctx):
    """at B = 600011 the c4 windows of the first grid rows exceed 2048, so the
    reference's 2048-wide unbounded tiles search part of them twice: with
    twin_ref_tiles = 2048 the counters and hits repeat exactly as the oracle's
    restatement of the reference loops says; without, every c4 once"""
    say("twin reference tiles 2048 at B=600011")
    B = 600011
    t = device_table(ctx, B // 42, 16)
    cands = O.twin_cands(B)[[97, 194, 679, 776]]  # candidates whose c3 sweeps are not all gated away
    ctx.set_option("twin_ref_tiles", 2048)
    try:
        res, hits, _ = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
    finally:
        ctx.set_option("twin_ref_tiles", 0)
    oc, oh, _ = O.twin_run(cands, t, threads=THREADS, tile=2048)
    compare(res, hits, oc, oh)
    res0, hits0, _ = ctx.twin_run(cands.view(fc5.TWIN_CAND_DTYPE))
    once, oh0, _ = O.twin_run(cands, t, threads=THREADS)
    compare(res0, hits0, once, oh0)
    assert oc["c4_total"] > once["c4_total"] and oc["calls"] > once["calls"]
