"""GPU: the exact sorted pair table (fc5_exact_*) -- the reference's sorted
pair index (QuadSumIndex, quad_sum.hpp:8-45) built and kept on the device.

Bar: the device entries equal the oracle's restatement entry for entry (the
oracle is pinned to the reference's own index by tests/test_exact.py); lookups
give exactly the pairs of the device pair search; with "verify_table" a run's
hits and solutions are those of the pair-search verification, and every
planted exact target comes back flagged FC5_HIT_EXTRA."""
import hashlib
import json
import os
import random
import sys

import numpy as np
import pytest

import fc5
import oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    yield c
    c.close()


def _entries(ctx):
    n, _, _ = ctx.exact_info()
    return ctx.exact_read(0, n)


@pytest.mark.parametrize("N", [1, 2, 40, 300])
def test_table_equals_oracle(ctx, N):
    say(f"exact table N={N}")
    ctx.exact_build(N)
    n, NN, s = ctx.exact_info()
    ent, so = O.exact_table(N)
    assert (n, NN, s) == (len(ent), N, so)
    assert _entries(ctx) == ent


def test_table_n1000_matches_reference_golden(ctx):
    say("exact table N=1000 vs the reference index golden")
    rec = next(r for r in json.load(open(os.path.join(GOLDEN, "exact_table.json"))) if r["N"] == 1000)
    ctx.exact_build(1000)
    _, _, s = ctx.exact_info()
    lines = [f"{v} {a} {b}" for v, a, b in (O.exact_decode(e, s) for e in _entries(ctx))]
    assert len(lines) == rec["count"]
    assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == rec["sha256"]


def test_table_with_extra_values_equals_oracle(ctx):
    say("exact table N=200 with extra values")
    rnd = random.Random(11)
    N = 200
    extra = [2 ** 6 + 3 ** 6, 5, 2 * N ** 6, 2 * N ** 6 + 1] + [rnd.randrange(1, 1 << 60) for _ in range(40)]
    ctx.exact_build(N, extra)
    ent, _ = O.exact_table(N, extra)
    assert _entries(ctx) == ent


def test_lookup_matches_pair_search(ctx):
    """random pair sums (with bounds below and above b), near misses and extra
    values: the lookup's pairs are the pair search's, extras come back (i, 0, 0)"""
    say("exact lookup vs pair search, N=2428")
    rnd = random.Random(5)
    N = 2428
    extra = [rnd.randrange(1, 2 * N ** 6) for _ in range(16)]
    ctx.exact_build(N, extra)
    targets, hi = [], []
    for _ in range(300):
        a = rnd.randint(1, N)
        b = rnd.randint(a, N)
        targets.append(a ** 6 + b ** 6 + rnd.choice((0, 0, 0, 1, -1)))
        hi.append(rnd.choice((N, b, max(1, b - 1), a)))
    targets += extra
    hi += [N] * len(extra)
    got = ctx.exact_lookup(targets, hi)
    want = ctx.pair_search(targets, hi)
    assert sorted(t for t in got if t[1]) == sorted(want)
    assert sorted(t[0] for t in got if not t[1]) == list(range(300, 300 + len(extra)))
    assert len(want) > 60  # about 300 * 3/5 exact sums * 1/2 bounds at or above b


def test_large_table_sorted_and_complete(ctx):
    """N = 24285 (the 1.0M band's table, 2.95e8 entries, 4.7 GB): strictly
    increasing, every entry a valid pair a <= b <= N (checked mod 2^64 with b
    from the float root), and exactly N(N+1)/2 of them -- so it is the set of
    all pairs; random pair sums are found by lookup"""
    N = 24285
    say(f"exact table N={N}: build")
    ctx.exact_build(N)
    n, _, s = ctx.exact_info()
    assert n == N * (N + 1) // 2
    say("read back")
    raw = np.zeros(2 * n, np.uint64)
    rc = fc5._lib.fc5_exact_read(ctx.h, 0, n, raw.ctypes.data_as(fc5._P))
    assert rc == 0
    lo, hi = raw[0::2], raw[1::2]
    say("sorted")
    inc = (hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))
    assert inc.all()
    del inc
    say("pairs")
    mask = np.uint64((1 << s) - 1)
    a = lo & mask
    assert a.min() >= 1 and a.max() <= N
    # value = entry >> s: low 64 bits of the value, and the value as a float
    v_lo = (lo >> np.uint64(s)) | (hi << np.uint64(64 - s))
    v_f = hi.astype(np.float64) * 2.0 ** (64 - s) + (lo >> np.uint64(s)).astype(np.float64)
    a6 = a * a * a
    a6 = a6 * a6
    b = np.rint(np.maximum(v_f - a.astype(np.float64) ** 6, 0.0) ** (1.0 / 6.0)).astype(np.uint64)
    ok = np.zeros(n, bool)
    for db in (0, 1, -1):
        bb = b + np.uint64(db) if db >= 0 else b - np.uint64(1)
        b6 = bb * bb * bb
        ok |= (a6 + b6 * b6 == v_lo) & (bb >= a) & (bb <= np.uint64(N))
    assert ok.all()
    del raw, lo, hi, v_lo, v_f, a6, b, ok
    rnd = random.Random(3)
    pairs = [(x, rnd.randint(x, N)) for x in (rnd.randint(1, N) for _ in range(200))]
    got = ctx.exact_lookup([x ** 6 + y ** 6 for x, y in pairs], [N] * len(pairs))
    assert sorted(got) == sorted((i, x, y) for i, (x, y) in enumerate(pairs))
    ctx.exact_clear()
    assert ctx.exact_info() == (0, 0, 0)


def test_band_verify_table_same_results_and_plants_flagged(ctx):
    """B in [100001,102000]: 64 planted keys in the xor table and their exact targets
    as extra values: verify by lookup = verify by pair search, plus the flag"""
    say("band run: verify_table vs pair search, 64 plants")
    _, Bs, us = fc5.emit_units(100001, 102000)
    pts = fc5.plant_points(Bs, us, 64, 0x615 << 8 | 2, exact=True)
    keys = np.array([p[3] for p in pts], np.uint64)
    ctx.table_build(2428, 48, extra=keys)
    units = fc5.units_array(Bs, us)
    ctx.set_option("verify_table", 0)
    ref = ctx.run(units)
    ctx.exact_build(2428, [p[4] for p in pts])
    ctx.set_option("verify_table", 1)
    try:
        out = ctx.run(units)
    finally:
        ctx.set_option("verify_table", 0)
    for k in ("calls", "gated", "probes", "maybe", "solutions"):
        assert out.result[k] == ref.result[k], k
    assert out.solutions == ref.solutions
    plants = {(i, e, d) for i, e, d, _, _ in pts}
    flagged = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits if h["flags"] & fc5.HIT_EXTRA}
    assert flagged == plants
    assert not any(int(h["flags"]) & fc5.HIT_EXTRA for h in ref.hits)
    stripped = out.hits.copy()
    stripped["flags"] &= ~np.uint32(fc5.HIT_EXTRA)
    assert np.array_equal(stripped, ref.hits)


def test_serving_with_verify_table(ctx):
    say("serving with verify_table")
    _, Bs, us = fc5.emit_units(100001, 102000)
    pts = fc5.plant_points(Bs, us, 16, 0x615 << 8 | 9, exact=True)
    ctx.table_build(2428, 48, extra=np.array([p[3] for p in pts], np.uint64))
    ctx.exact_build(2428, [p[4] for p in pts])
    units = fc5.units_array(Bs, us)
    ctx.set_option("verify_table", 1)
    try:
        alone = ctx.run(units)
        ctx.serve_start(max_bands=2, round_points=1e8)
        try:
            bid = ctx.serve_submit(units)
            assert ctx.serve_wait(bid, 120.0)
            got = ctx.serve_result(bid)
        finally:
            ctx.serve_stop()
    finally:
        ctx.set_option("verify_table", 0)
    assert np.array_equal(got.hits, alone.hits)
    assert sum(1 for h in got.hits if h["flags"] & fc5.HIT_EXTRA) == 16


def test_errors(ctx):
    say("exact table errors")
    ctx.table_build(2428, 8)
    ctx.exact_clear()
    _, Bs, us = fc5.emit_units(100001, 100400)
    units = fc5.units_array(Bs, us)
    ctx.set_option("verify_table", 1)
    try:
        # no exact table: the run's verification refuses (r = 8 gives many maybe-hits)
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.run(units)
        assert e.value.code == fc5.E_STATE
        ctx.exact_build(100)  # N below the pair bound of the hits
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.run(units)
        assert e.value.code == fc5.E_STATE
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.serve_start()
        assert e.value.code == fc5.E_STATE
    finally:
        ctx.set_option("verify_table", 0)
    for N, extra in ((0, ()), (262144, ()), (10, (0,)), (10, (1 << 125,))):
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.exact_build(N, extra)
        assert e.value.code == fc5.E_ARG, (N, extra)
    ctx.exact_build(10)
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.exact_read(50, 10)
    assert e.value.code == fc5.E_ARG
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.exact_lookup([65], [11])  # bound above the table's N
    assert e.value.code == fc5.E_STATE
    assert ctx.exact_lookup([65, 66, 2], [10, 10, 10]) == [(0, 1, 2), (2, 1, 1)]
    ctx.exact_clear()


def test_find3_verify_table_flags_planted_targets(ctx):
    """classes 2-4: planted keys (T(d) - c3^6) mod 2^64 in the xor table and their
    exact targets T(d) - c3^6 as extra values of the exact table; verifying the
    maybe-hits by lookup flags exactly the plants and leaves every other hit as
    the pair search has it"""
    say("find3 with verify_table")
    cls, Bs, us = fc5.emit_units(2000001, 2000400, classes=(2, 3, 4))
    rng = np.random.default_rng(12)
    pts = []
    for i in rng.choice(len(Bs), 16, replace=False):
        k, B, u = int(cls[i]), int(Bs[i]), int(us[i])
        res, mod, fmax = O.free_d_cls234(B, u, k)
        T = None
        for _ in range(200):
            d = int(res[rng.integers(len(res))] + mod * rng.integers(fmax // mod + 1))
            T = O.T_peel1(B, u, {2: 14, 3: 21, 4: 7}[k], d) if 1 <= d <= fmax else None
            if T:
                break
        if not T:
            continue
        hi = min(int(round(T ** (1 / 6))) + 2, (B - 1) // 42)
        while hi ** 6 > T:
            hi -= 1
        lo = 1
        while lo <= hi and 3 * lo ** 6 < T:
            lo += 1
        if lo > hi:
            continue
        c3 = int(rng.integers(lo, hi + 1))
        pts.append((int(i), d, c3, T - c3 ** 6))
    assert len(pts) >= 8
    ctx.table_build(2428, 48, extra=np.array([v % 2 ** 64 for *_, v in pts], np.uint64))
    units = fc5.units_array(Bs, us)
    ctx.set_option("use_gate", 0)  # plants need not pass the residue gate
    try:
        ctx.set_option("verify_table", 0)
        ref = ctx.run(units, cls=cls)
        ctx.exact_build(2428, [v for *_, v in pts])
        ctx.set_option("verify_table", 1)
        out = ctx.run(units, cls=cls)
    finally:
        ctx.set_option("verify_table", 0)
        ctx.set_option("use_gate", 1)
        ctx.exact_clear()
    flagged = {(int(h["unit"]), int(h["d"]), int(h["e"])) for h in out.hits if h["flags"] & fc5.HIT_EXTRA}
    assert flagged == {(i, d, c) for i, d, c, _ in pts}
    stripped = out.hits.copy()
    stripped["flags"] &= ~np.uint32(fc5.HIT_EXTRA)
    assert np.array_equal(stripped, ref.hits)
    assert out.solutions == ref.solutions


def test_twin_verify_table_finds_the_known(ctx):
    """the 54 known (6,1,7) class-5 solutions, verified by lookups of
    T - c4^6 - c3^6 in the exact sorted pair table instead of the pair search"""
    import json
    say("twin with verify_table")
    known = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_617_cls5.json")))
    N = max(k["B"] for k in known) // 42
    ctx.table_build(N, 48, on_gpu=True)
    rows = []
    for k in known:
        c = fc5.twin_cands(k["B"])
        rows.append(c[(c["u"] == k["u"]) & (c["w"] == k["w"])])
    cands = np.concatenate(rows)
    res0, hits0, sols0 = ctx.twin_run(cands)
    ctx.exact_build(N)
    ctx.set_option("verify_table", 1)
    try:
        res1, hits1, sols1 = ctx.twin_run(cands)
    finally:
        ctx.set_option("verify_table", 0)
        ctx.exact_clear()
    assert sols1 == sols0
    assert sorted(tuple(s["terms"]) for s in sols1 if s["gcd"] == 1) == sorted(tuple(k["terms"]) for k in known)
    assert np.array_equal(hits1, hits0)
