"""GPU: fc5_run_units_noalloc (no device allocation; its buffers live in a
static device array per GPU) against fc5_run_units on the same inputs and the
same table bytes: counters, the sorted maybe-hit list (r = 16, so it is not
empty) and the solutions must be identical, on two calls in a row, the second
with a smaller input (the array keeps the first call's contents)."""
import sys

import numpy as np
import pytest

import fc5

pytestmark = pytest.mark.gpu

KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe", "n_hits", "hits_dropped", "solutions",
        "units", "units_skipped")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    c.table_build(2428, 16)
    yield c
    c.close()


def _same(alloc, free):
    for k in KEYS:
        assert free.result[k] == alloc.result[k], (k, free.result[k], alloc.result[k])
    assert np.array_equal(free.hits, alloc.hits)
    assert free.solutions == alloc.solutions
    assert free.rc == fc5.OK


def test_noalloc_matches_run_units_twice_smaller_second(ctx):
    _, Bs, us = fc5.emit_units(100001, 102000)
    big = fc5.units_array(Bs, us)
    _, Bs2, us2 = fc5.emit_units(3200001, 3200400)
    small = fc5.units_array(Bs2, us2)[::7]  # past the 2^128 wall, fewer units
    assert len(small) < len(big)
    for units in (big, small):
        say(f"noalloc: {len(units)} units")
        alloc = ctx.run(units, fc5.MODE_BATCHED)
        free = ctx.run_noalloc(units)
        assert alloc.result["n_hits"] > 0
        _same(alloc, free)


def test_noalloc_limits_refused(ctx):
    too_many = fc5.units_array([100003] * (fc5.NOALLOC_UNITS + 1), [17646] * (fc5.NOALLOC_UNITS + 1))
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.run_noalloc(too_many)
    assert e.value.code == fc5.E_ARG
    _, Bs, us = fc5.emit_units(fc5.NOALLOC_BMAX + 1, fc5.NOALLOC_BMAX + 200)
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.run_noalloc(fc5.units_array(Bs, us))
    assert e.value.code == fc5.E_ARG


def test_noalloc_hit_overflow_reported(ctx):
    """more maybe-hits than the array's hit buffer holds: FC5_E_OVERFLOW, the kept
    hits are a subset of fc5_run_units' hits, and the counters are unaffected"""
    c = fc5.Context(0)
    try:
        c.table_build(2428, 4)  # 4-bit cells: one false positive in 16 probes
        c.set_option("use_gate", 0)
        c.set_option("verify", 0)
        _, Bs, us = fc5.emit_units(100001, 102000)
        units = fc5.units_array(Bs, us)
        alloc = c.run(units, fc5.MODE_BATCHED)
        total = alloc.result["n_hits"]
        assert total > fc5.NOALLOC_HITS and alloc.result["hits_dropped"] == 0
        free = c.run_noalloc(units, hit_cap=fc5.NOALLOC_HITS, allow_overflow=True)
        assert free.rc == fc5.E_OVERFLOW
        assert free.result["n_hits"] == fc5.NOALLOC_HITS == len(free.hits)
        assert free.result["hits_dropped"] == total - fc5.NOALLOC_HITS
        for k in ("calls", "neg", "zero", "gated", "probes", "maybe"):
            assert free.result[k] == alloc.result[k], k
        key = lambda h: (int(h["unit"]), int(h["e"]), int(h["d"]), int(h["flags"]), int(h["fp"]))
        assert {key(h) for h in free.hits} <= {key(h) for h in alloc.hits}
    finally:
        c.close()
