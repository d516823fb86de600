"""GPU parity of the strict residue gate (option "strict_gate") against the oracle.

With the option on, every kernel form (batched, chunked, plain; classes 2-5)
must give the oracle's strict counters (calls, neg, zero, gated, probes equal
to the default run's; maybe and side_maybe over the strict survivors;
strict_out = probes - queries) and its maybe-hit list bit for bit, on the same
table bytes.  The oracle's strict mode is pinned by the big-integer brute force
in test_strict_gate.py.  The hit list must also be the default run's hit list
restricted to the hits whose exact T passes the strict gate.
"""
import os
import sys

import numpy as np
import pytest

import fc5
import oracle as O

pytestmark = pytest.mark.gpu

C5 = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe")
F3 = ("jobs", "fail_t", "calls", "gated", "probes", "maybe")
THREADS = min(16, os.cpu_count() or 8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    yield c
    c.close()


@pytest.fixture
def strict(ctx):
    ctx.set_option("strict_gate", 1)
    yield ctx
    ctx.set_option("strict_gate", 0)


def _load(ctx, N, r, extra=None):
    ctx.table_build(N, r, extra=extra)
    t = O.Table.build_pairs(N, r, extra=extra)
    h = ctx.table_info()
    assert (h.n_keys, h.m_cells, h.mix_seed) == (t.x.hdr.n_keys, t.x.hdr.m_cells, t.x.hdr.mix_seed)
    return t


def _oracle_mixed(cls, Bs, us, table):
    """per-class strict oracle counters (with strict_out) and the hit list in caller unit indices"""
    out, hits = {}, []
    idx5 = np.nonzero(cls == 5)[0]
    if len(idx5):
        arr, ok = O.pack_units(Bs[idx5], us[idx5])
        assert ok.all()
        c, h, _ = O.peel(arr, len(idx5), table, threads=THREADS, strict=True)
        c["strict_out"] = c["probes"] - c["queries"]
        out[5] = c
        h = h.copy()
        h["unit"] = idx5[h["unit"]]
        hits.append(h)
    for k in (2, 3, 4):
        sel = np.nonzero(cls == k)[0]
        if not len(sel):
            continue
        c, h, _ = O.find3(cls[sel], Bs[sel], us[sel], table, threads=THREADS, strict=True)
        out[k] = {"jobs": c["jobs"], "fail_t": c["failT"], "calls": c["calls"], "gated": c["gated"],
                  "probes": c["probes"], "maybe": c["maybe"], "strict_out": c["probes"] - c["queries"]}
        h = h.copy()
        h["unit"] = sel[h["unit"]]
        hits.append(h)
    allh = np.concatenate(hits) if hits else np.zeros(0, O.HIT_DTYPE)
    return out, np.sort(allh, order=["unit", "e", "d", "flags"])


def _compare(out, oc, oh):
    for k, ref in oc.items():
        got = out.by_class[k]
        for key in (C5 if k == 5 else F3) + ("strict_out",):
            assert got[key] == ref[key], (k, key, got[key], ref[key])
    got = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(got) == len(oh)
    for f in ("unit", "d", "e", "flags", "fp"):
        assert np.array_equal(got[f], oh[f]), f


def _hit_T(cls, B, u, d, e):
    if cls == 5:
        return (B ** 6 - u ** 6 - (14 * e) ** 6 - (21 * d) ** 6) // 42 ** 6
    f = {2: 14, 3: 21, 4: 7}[cls]
    return (B ** 6 - u ** 6 - (f * d) ** 6) // 42 ** 6 - e ** 6


def _filtered(default_hits, cls, Bs, us):
    keep = [h for h in default_hits
            if O.strict_pass_T(_hit_T(int(cls[h["unit"]]), int(Bs[h["unit"]]), int(us[h["unit"]]),
                                      int(h["d"]), int(h["e"])))]
    return sorted((int(h["unit"]), int(h["e"]), int(h["d"]), int(h["fp"])) for h in keep)


def _keyset(hits):
    return sorted((int(h["unit"]), int(h["e"]), int(h["d"]), int(h["fp"])) for h in hits)


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_CHUNKED, fc5.MODE_PLAIN],
                         ids=["batched", "chunked", "plain"])
def test_strict_band_100001_102000_matches_oracle(ctx, strict, mode):
    say(f"strict gate, B in [100001,102000], r=8, mode={mode}")
    cls, Bs, us = fc5.emit_units(100001, 102000)
    t = _load(ctx, 2428, 8)  # r=8: false positives frequent enough to pass the strict gate
    out = ctx.run(fc5.units_array(Bs, us), mode)
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare(out, oc, oh)
    assert len(oh) > 0 and 0 < out.result["strict_out"] < out.result["probes"]
    ctx.set_option("strict_gate", 0)
    base = ctx.run(fc5.units_array(Bs, us), mode)
    ctx.set_option("strict_gate", 1)
    assert base.result["strict_out"] == 0 and base.result["probes"] == out.result["probes"]
    assert _filtered(base.hits, cls, Bs, us) == _keyset(out.hits) and len(base.hits) > len(out.hits)


@pytest.mark.parametrize("lo,hi,every", [(1000001, 1020000, 97), (3200001, 3210000, 997)])
def test_strict_sampled_units_match_oracle(ctx, strict, lo, hi, every):
    """unit samples of the 1.0M and 3.2M bands (B^6 > 2^128 in the second), batched and plain"""
    say(f"strict gate, sampled units [{lo},{hi}] every {every}")
    cls, Bs, us = fc5.emit_units(lo, hi)
    cls, Bs, us = cls[::every], Bs[::every], us[::every]
    t = _load(ctx, 2428, 8)  # r=8: false positives frequent enough to pass the strict gate
    b = ctx.band(fc5.units_array(Bs, us))
    ob = b.run(fc5.MODE_BATCHED)
    op = b.run(fc5.MODE_PLAIN)
    b.close()
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare(ob, oc, oh)
    _compare(op, oc, oh)
    assert len(oh) > 0


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN], ids=["batched", "plain"])
def test_strict_mixed_classes_match_oracle(ctx, strict, mode):
    """classes 2-5 of B in [100001,102000] in one band"""
    say(f"strict gate, mixed classes B in [100001,102000], mode={mode}")
    cls, Bs, us = fc5.emit_units(100001, 102000, classes=(2, 3, 4, 5))
    t = _load(ctx, 2428, 8)  # r=8: false positives frequent enough to pass the strict gate
    out = ctx.run(fc5.units_array(Bs, us), mode, cls=cls)
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare(out, oc, oh)
    assert all(out.by_class[k]["strict_out"] > 0 for k in (2, 3, 4, 5))
    # classes 2-4 make queries too (their false positives mostly fall to the strict gate here)
    assert sum(out.by_class[k]["probes"] - out.by_class[k]["strict_out"] for k in (2, 3, 4)) > 0


def test_strict_mixed_2000001_sample_and_ragged_lines(ctx, strict):
    """a 2.0M sample of all four classes, plus .buc lines the batched kernel
    hands to the per-point 192-bit path (irregular units)"""
    say("strict gate, mixed 2.0M sample and ragged lines")
    cls, Bs, us = fc5.emit_units(2000001, 2010000, classes=(2, 3, 4, 5))
    sel = np.concatenate([np.nonzero(cls == k)[0][::m] for k, m in ((2, 60), (3, 60), (4, 400), (5, 400))])
    sel.sort()
    cls, Bs, us = cls[sel], Bs[sel], us[sel]
    t = _load(ctx, 2428, 8)  # r=8: false positives frequent enough to pass the strict gate
    out = ctx.run(fc5.units_array(Bs, us), fc5.MODE_BATCHED, cls=cls)
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare(out, oc, oh)
    # irregular lines: even u that is no seed lift, at 0.1M and 3.2M
    rB, rU = [], []
    for lo in (100001, 3200001):
        _, B1, _ = fc5.emit_units(lo, lo + 60)
        for b in sorted({int(x) for x in B1})[:3]:
            for k in range(3):
                rB.append(b)
                rU.append(2 * (b // 4 + 977 * k + 1))
    rB, rU = np.array(rB, np.uint64), np.array(rU, np.uint64)
    arr, ok = O.pack_units(rB, rU)
    keep = np.flatnonzero(ok)
    res = ctx.run(fc5.units_array(rB, rU), fc5.MODE_BATCHED)
    assert res.result["units_irregular"] > 0
    sub = (O.Unit5 * len(keep))(*[arr[int(i)] for i in keep])
    c, h, _ = O.peel(sub, len(keep), t, threads=THREADS, strict=True)
    for k in C5:
        assert res.result[k] == c[k], k
    assert res.result["strict_out"] == c["probes"] - c["queries"]
    h = h.copy()
    h["unit"] = keep[h["unit"]].astype(np.uint32)
    assert _keyset(res.hits) == _keyset(h)


def test_strict_plants_graph_and_tile_shapes(ctx, strict):
    """plants that pass the strict gate are all found; graph replays and every
    tile shape give the eager result"""
    say("strict gate, plants, graph replays, tile shapes")
    cls, Bs, us = fc5.emit_units(1000001, 1004000)
    pts = fc5.plant_points(Bs, us, 64, 0x615 << 8 | 11, strict=True)
    t = _load(ctx, 2428, 16)
    ctx.side_build(np.array([p[3] for p in pts], np.uint64), 48)
    try:
        b = ctx.band(fc5.units_array(Bs, us))
        ref = b.run(fc5.MODE_BATCHED)
        found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in ref.hits if h["flags"] & 2}
        assert {(p[0], p[1], p[2]) for p in pts} <= found
        ctx.set_option("use_graph", 1)
        for _ in range(3):
            g = b.run(fc5.MODE_BATCHED)
            assert g.result == {**ref.result, "kernel_ms": g.result["kernel_ms"], "wall_ms": g.result["wall_ms"]}
            assert _keyset(g.hits) == _keyset(ref.hits)
        ctx.set_option("use_graph", 0)
        for rows in (4, 8, 16, 32):
            ctx.set_option("rows_per_item", rows)
            o = b.run(fc5.MODE_BATCHED)
            for k in C5 + ("strict_out",):
                assert o.result[k] == ref.result[k], (rows, k)
            assert _keyset(o.hits) == _keyset(ref.hits)
        ctx.set_option("rows_per_item", 0)
        b.close()
        # the main-table part against the oracle (the side table is the product's own build)
        oc, oh = _oracle_mixed(cls, Bs, us, t)
        for k in ("calls", "neg", "zero", "gated", "probes", "maybe", "strict_out"):
            assert ref.result[k] == oc[5][k], k
        assert _keyset(ref.hits[(ref.hits["flags"] & 1) != 0]) == _keyset(oh)
    finally:
        ctx.set_option("use_graph", 0)
        ctx.set_option("rows_per_item", 0)
        ctx.side_clear()
