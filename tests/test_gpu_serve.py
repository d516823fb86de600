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
        # bands resident at the same time always fit the arenas together
        for j in range(6):
            live = [k for k in range(6) if t[k]["t_admit"] <= t[j]["t_admit"] < t[k]["t_done"]]
            assert sum(size5[k] for k in live) <= n5 + 16 and sum(size3[k] for k in live) <= n3 + 16
        assert max(x["t_admit"] for x in t) > min(x["t_done"] for x in t)  # some band waited for room
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


def test_serving_hit_overflow_is_flagged(ctx):
    """a round whose hits exceed hit_cap: the band finishes with FC5_E_OVERFLOW,
    hits_dropped = 1 (the round's buffer is shared, so the count is not split
    among bands), the kept hits are a subset of the band's hits alone, and the
    counters are unaffected"""
    say("serving: hit overflow in a round")
    cls, Bs, us = fc5.emit_units(1000001, 1002000, classes=(4, 5))
    units = fc5.units_array(Bs, us)
    alone = ctx.run(units, fc5.MODE_BATCHED, cls=cls)
    n = alone.result["n_hits"]
    assert n >= 2
    ctx.set_option("hit_cap", n - 1)
    try:
        ctx.serve_start(round_points=1e15)  # the whole band in one round
        try:
            bid = ctx.serve_submit(units, cls)
            assert ctx.serve_wait(bid, 120.0)
            with pytest.raises(fc5.Fc5Error) as e:
                ctx.serve_result(bid)
            assert e.value.code == fc5.E_OVERFLOW
            out = ctx.serve_result(bid, allow_overflow=True)
            assert out.rc == fc5.E_OVERFLOW
            assert out.result["hits_dropped"] == 1
            assert out.result["n_hits"] == len(out.hits) == n - 1
            key = lambda h: (int(h["unit"]), int(h["e"]), int(h["d"]), int(h["flags"]))
            assert {key(h) for h in out.hits} <= {key(h) for h in alone.hits}
            for f in ("calls", "neg", "zero", "gated", "probes", "maybe", "jobs", "fail_t"):
                assert out.result[f] == alone.result[f], f
        finally:
            ctx.serve_stop()
    finally:
        ctx.set_option("hit_cap", 1 << 20)
