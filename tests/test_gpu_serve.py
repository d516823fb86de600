"""GPU: the resident serving loop (fc5_serve_*): continuous batching of bands
that arrive while others run, admission by counter slots and arena space, and
results identical to running each band alone (fc5_band_run)."""
import math
import random
import sys
import time

import numpy as np
import pytest

import fc5

pytestmark = pytest.mark.gpu

KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe", "jobs", "fail_t", "units")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_bands(seed, n, lo, hi, wmin, wmax, classes=(4, 5)):
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        w = int(math.exp(rnd.uniform(math.log(wmin), math.log(wmax))))
        b0 = rnd.randint(lo, hi)
        cls, Bs, us = fc5.emit_units(b0, b0 + w - 1, classes=classes)
        out.append((cls, fc5.units_array(Bs, us)))
    return out


def same(alone, served):
    for k in (2, 3, 4, 5):
        for f in KEYS:
            assert served.by_class[k][f] == alone.by_class[k][f], (k, f)
    for f in ("calls", "probes", "maybe", "n_hits", "solutions", "units", "units_skipped"):
        assert served.result[f] == alone.result[f], f
    assert np.array_equal(served.hits, alone.hits)
    assert served.solutions == alone.solutions


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    c.table_build(1206000 // 42, 16, on_gpu=True)  # r = 16: non-empty hit lists
    yield c
    c.close()


def test_serving_matches_alone_with_overlap(ctx):
    say("serving: 10 bands, staggered, results alone vs mixed")
    bands = make_bands(615, 10, 1000000, 1200000, 200, 3000)
    alone = [ctx.run(u, fc5.MODE_BATCHED, cls=c) for c, u in bands]
    assert sum(a.result["maybe"] for a in alone) > 0
    ctx.serve_start(max_bands=4, round_points=1e9)
    try:
        ids = [ctx.serve_submit(u, c) for c, u in bands[:5]]
        time.sleep(0.05)
        ids += [ctx.serve_submit(u, c) for c, u in bands[5:]]
        for i in ids:
            assert ctx.serve_wait(i, 120.0)
        for i, a in zip(ids, alone):
            same(a, ctx.serve_result(i))
        t = [ctx.serve_timing(i) for i in ids]
        st = ctx.serve_stats()
        assert st["done"] == 10 and st["rounds"] > 10
        # some band was admitted while an earlier one was still running
        assert any(t[j]["t_admit"] < t[k]["t_done"] and t[j]["t_done"] > t[k]["t_admit"]
                   for j in range(10) for k in range(j))
        assert all(x["rounds"] >= 1 for x in t)
        with pytest.raises(fc5.Fc5Error) as e:  # the context is serving
            ctx.run(bands[0][1], fc5.MODE_BATCHED, cls=bands[0][0])
        assert e.value.code == fc5.E_STATE
        for i in ids:
            ctx.serve_release(i)
    finally:
        ctx.serve_stop()


def test_serving_admission_queues_by_arena_space(ctx):
    say("serving: arena admission")
    bands = make_bands(616, 6, 1000000, 1200000, 500, 2500)
    alone = [ctx.run(u, fc5.MODE_BATCHED, cls=c) for c, u in bands]
    n5 = max(int((c == 5).sum()) for c, _ in bands)
    n3 = max(int((c != 5).sum()) for c, _ in bands)
    ctx.serve_start(max_bands=8, arena_units5=n5 + 16, arena_units3=n3 + 16, round_points=2e9)
    try:
        ids = [ctx.serve_submit(u, c) for c, u in bands]
        st = ctx.serve_stats()
        assert st["submitted"] == 6
        for i in ids:
            assert ctx.serve_wait(i, 120.0)
        for i, a in zip(ids, alone):
            same(a, ctx.serve_result(i))
        t = [ctx.serve_timing(i) for i in ids]
        size5 = [int((c == 5).sum()) for c, _ in bands]
        size3 = [int((c != 5).sum()) for c, _ in bands]
        # bands resident at the same time always fit the arenas together (a band holds its
        # arena ranges from admission until its last round has ended)
        for j in range(6):
            live = [k for k in range(6) if t[k]["t_admit"] <= t[j]["t_admit"] < t[k]["t_release"]]
            assert sum(size5[k] for k in live) <= n5 + 16 and sum(size3[k] for k in live) <= n3 + 16
        assert all(x["t_admit"] <= x["t_release"] <= x["t_done"] for x in t)
        assert max(x["t_admit"] for x in t) >= min(x["t_release"] for x in t)  # some band waited for room
    finally:
        ctx.serve_stop()


def test_serving_ranges_emitted_on_device(ctx):
    """fc5_serve_submit_range: Stage-1 on a side stream while the server runs"""
    say("serving: submitted B ranges")
    ranges = [(1000001, 1001500), (1100001, 1100900), (1150001, 1152500), (1199001, 1200000)]
    alone = []
    for lo, hi in ranges:
        c, Bs, us = fc5.emit_units(lo, hi, classes=(4, 5))
        alone.append(ctx.run(fc5.units_array(Bs, us), fc5.MODE_BATCHED, cls=c))
    ctx.serve_start(max_bands=3, round_points=1e9)
    try:
        ids = [ctx.serve_submit_range(lo, hi, classes=(4, 5)) for lo, hi in ranges]
        for i in ids:
            assert ctx.serve_wait(i, 120.0)
        for i, a in zip(ids, alone):
            same(a, ctx.serve_result(i))
    finally:
        ctx.serve_stop()


def test_serving_refusals(ctx):
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.serve_submit(fc5.units_array([1000003], [6]))
    assert e.value.code == fc5.E_STATE
    ctx.serve_start()
    try:
        cls, Bs, us = fc5.emit_units(3000001, 3000100, classes=(5,))
        with pytest.raises(fc5.Fc5Error) as e:  # beyond what the table's N covers
            ctx.serve_submit(fc5.units_array(Bs, us), cls)
        assert e.value.code == fc5.E_ARG
        with pytest.raises(fc5.Fc5Error):
            ctx.serve_result(12345)
        empty = ctx.serve_submit(fc5.units_array([], []))
        assert ctx.serve_wait(empty, 30.0)
        assert ctx.serve_result(empty).result["units"] == 0
    finally:
        ctx.serve_stop()


def test_serving_hit_overflow_is_charged_per_band(ctx):
    """rounds whose hits exceed hit_cap: each band is charged exactly its own
    dropped hits (n_hits + hits_dropped equals its hit count alone), returns
    FC5_E_OVERFLOW only when some of its hits were dropped, keeps a subset of
    its hits alone, and its counters are unaffected"""
    say("serving: hit overflow charged per band")
    ranges = [(1000001, 1002000), (1100001, 1101000), (1150001, 1150600)]
    bands = []
    for lo, hi in ranges:
        cls, Bs, us = fc5.emit_units(lo, hi, classes=(4, 5))
        units = fc5.units_array(Bs, us)
        bands.append((cls, units, ctx.run(units, fc5.MODE_BATCHED, cls=cls)))
    total = sum(a.result["n_hits"] for *_, a in bands)
    assert min(a.result["n_hits"] for *_, a in bands) >= 2
    cap = total // 3  # below the largest band's hit count
    ctx.set_option("hit_cap", cap)
    try:
        ctx.serve_start(max_bands=4, round_points=1e15)  # a band's items in one round
        try:
            ids = [ctx.serve_submit(u, c) for c, u, _ in bands]
            for i in ids:
                assert ctx.serve_wait(i, 120.0)
            dropped = 0
            for i, (_, _, alone) in zip(ids, bands):
                out = ctx.serve_result(i, allow_overflow=True)
                n = alone.result["n_hits"]
                assert out.result["n_hits"] == len(out.hits)
                assert out.result["n_hits"] + out.result["hits_dropped"] == n
                assert (out.rc == fc5.E_OVERFLOW) == (out.result["hits_dropped"] > 0)
                key = lambda h: (int(h["unit"]), int(h["e"]), int(h["d"]), int(h["flags"]))
                assert {key(h) for h in out.hits} <= {key(h) for h in alone.hits}
                if out.result["hits_dropped"] == 0:
                    assert np.array_equal(out.hits, alone.hits)
                for f in ("calls", "neg", "zero", "gated", "probes", "maybe", "jobs", "fail_t"):
                    assert out.result[f] == alone.result[f], f
                dropped += out.result["hits_dropped"]
            # however the bands were grouped into rounds, a band's hits beyond the cap are dropped
            assert dropped >= max(a.result["n_hits"] for *_, a in bands) - cap > 0
        finally:
            ctx.serve_stop()
    finally:
        ctx.set_option("hit_cap", 1 << 20)


def test_serving_with_the_strict_gate_matches_alone(ctx):
    """the option strict_gate on a served mix: every band's counters (strict_out
    included) and hits are its result alone, and its hit list is the default
    one restricted to the keys that pass the strict gate"""
    say("serving: strict gate, 6 bands alone vs mixed")
    bands = make_bands(6150, 6, 1000000, 1200000, 200, 2000)
    base = [ctx.run(u, fc5.MODE_BATCHED, cls=c) for c, u in bands]
    ctx.set_option("strict_gate", 1)
    try:
        alone = [ctx.run(u, fc5.MODE_BATCHED, cls=c) for c, u in bands]
        ctx.serve_start(max_bands=3, round_points=1e9)
        try:
            ids = [ctx.serve_submit(u, c) for c, u in bands]
            for i in ids:
                assert ctx.serve_wait(i, 120.0)
            for i, a, b0 in zip(ids, alone, base):
                r = ctx.serve_result(i)
                same(a, r)
                assert r.result["strict_out"] == a.result["strict_out"] > 0
                assert r.result["probes"] == b0.result["probes"] and r.result["n_hits"] <= b0.result["n_hits"]
                assert set(r.hits.tolist()) <= set(b0.hits.tolist())
                ctx.serve_release(i)
        finally:
            ctx.serve_stop()
    finally:
        ctx.set_option("strict_gate", 0)
