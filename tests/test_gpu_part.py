"""GPU: partitioned probing (probe_mode 1) against gathers and the oracle.

The partitioned run must give the same per-class counters and the same
maybe-hit records (unit, e, d, fp, flags: main and side table) as the run
with three gathers per probe (probe_mode 0), which the other GPU tests pin
to the oracle; one band is also compared with the oracle directly.  Small
emit buffers force many pieces and the direct-query fallback of probes that
find no free chunk; r = 16 and r = 24 tables make the accepted sets and the
8-bit candidate stage non-trivial.
"""
import os
import sys

import numpy as np
import pytest

import fc5
import oracle as O

pytestmark = pytest.mark.gpu

C5_KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe")
F3_KEYS = ("jobs", "fail_t", "calls", "gated", "probes", "maybe", "side_maybe")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    c.set_option("verify", 0)  # the maybe-hit records are compared; their exact check is tested elsewhere
    yield c
    c.close()


def _run(ctx, units, cls=None, probe_mode=0, records=None):
    ctx.set_option("probe_mode", probe_mode)
    if records:
        ctx.set_option("part_records", records)
    try:
        return ctx.run(units, fc5.MODE_BATCHED, cls=cls)
    finally:
        ctx.set_option("probe_mode", -1)
        ctx.set_option("part_records", (1 << 32) - 4096)


def _same(a, b, classes=(5,)):
    for k in classes:
        keys = C5_KEYS if k == 5 else F3_KEYS
        ra = a.by_class[k] if a.by_class else a.result
        rb = b.by_class[k] if b.by_class else b.result
        for key in keys:
            assert ra[key] == rb[key], (k, key, ra[key], rb[key])
    ga = np.sort(a.hits, order=["unit", "e", "d", "flags"])
    gb = np.sort(b.hits, order=["unit", "e", "d", "flags"])
    assert len(ga) == len(gb)
    for f in ("unit", "d", "e", "fp", "flags"):
        assert np.array_equal(ga[f], gb[f]), f


@pytest.mark.parametrize("N,r", [(6000, 16), (6000, 24)])
def test_measure_partitioned_equals_direct_queries(ctx, N, r):
    ctx.table_build(N, r, on_gpu=True)
    ms, c = ctx.measure_partitioned(1 << 24)
    say(f"part/measure N={N} r={r}: {c}")
    assert c["lost"] == 0
    assert c["accepted"] == c["direct_accepted"]
    assert c["accepted_hsum"] == c["direct_hsum"]
    # r = 16: about 2^8 accepted keys; r = 24: about 2^16 8-bit candidates, each fully queried
    assert (c["accepted"] if r == 16 else c["full_queries"]) > 100


@pytest.fixture(scope="module")
def band_1m(ctx):
    """class-5 units of B in [1000001,1000400], an r = 16 device table (so the
    maybe-hit list is long) and the oracle's view of the same bytes, and the
    band searched with gathers"""
    _, Bs, us = fc5.emit_units(1000001, 1000400)
    units = fc5.units_array(Bs, us)
    ctx.table_build(23810, 16, on_gpu=True)
    h, packed = ctx.table_download()
    table = O.Table.from_bytes({f: getattr(h, f) for f, _ in fc5.StoreHeader._fields_}, packed)
    return Bs, us, units, table, _run(ctx, units, probe_mode=0)


@pytest.mark.parametrize("records", [1 << 16, 1 << 22], ids=["tiny-buffer", "one-piece"])
def test_class5_band_partitioned_equals_gathers_and_oracle(ctx, band_1m, records):
    Bs, us, units, t, direct = band_1m
    part = _run(ctx, units, probe_mode=1, records=records)
    st = ctx.part_stats()
    say(f"part/cls5 records={records}: {st[1]}, maybe={part.result['maybe']}")
    assert part.result["maybe"] > 0
    _same(direct, part)
    if records == 1 << 16:
        arr, ok = O.pack_units(Bs, us)
        oc, oh, _ = O.peel(arr, len(Bs), t, threads=min(16, os.cpu_count() or 8))
        for k in C5_KEYS[:6]:
            assert part.result[k] == oc[k], k
        got = np.sort(part.hits, order=["unit", "e", "d", "flags"])
        assert len(got) == len(oh)
        for f in ("unit", "d", "e", "fp"):
            assert np.array_equal(got[f], oh[f]), f


def test_mixed_band_with_side_table_partitioned_equals_gathers(ctx):
    """classes 2-5 over an r = 24 table, planted keys in the side table"""
    cls, Bs, us = fc5.emit_units(2000001, 2000100, classes=(2, 3, 4, 5))
    units = fc5.units_array(Bs, us)
    ctx.table_build(47620, 24, on_gpu=True)
    i5 = np.flatnonzero(cls == 5)
    pts = fc5.plant_points(Bs[i5], us[i5], 48, 0x615 << 8 | 9)
    ctx.side_build(np.array([k for *_, k in pts], np.uint64))
    try:
        direct = _run(ctx, units, cls=cls, probe_mode=0)
        part = _run(ctx, units, cls=cls, probe_mode=1, records=1 << 20)
    finally:
        ctx.side_clear()
    st = ctx.part_stats()
    say(f"part/mixed: {st}, side_maybe={part.by_class[5]['side_maybe']}")
    assert part.by_class[5]["side_maybe"] >= 48
    _same(direct, part, classes=(2, 3, 4, 5))


def test_headline_sample_partitioned_equals_gathers(ctx):
    """every 97th unit of the 3.2M band (B^6 > 2^128) over the N = 76,428, r = 48
    table: the default mode there is partitioned"""
    _, Bs, us = fc5.emit_units(3200001, 3210000)
    units = fc5.units_array(Bs[::97], us[::97])
    ctx.table_build(76428, 48, on_gpu=True)
    part = ctx.run(units, fc5.MODE_BATCHED)  # probe_mode -1: partitioned for this table
    st = ctx.part_stats()
    direct = _run(ctx, units, probe_mode=0)
    say(f"part/3.2M sample: {st}, probes={part.result['probes']}, maybe={part.result['maybe']}")
    assert st[1]["chunks"] > 0
    _same(direct, part)


def test_flagship_table_512_coarse_buckets_partitioned_equals_gathers(ctx):
    """the reference's flagship table, N = 119,047 (8.7e9 cells: 346 coarse
    buckets per third, so the 512-bucket kernel variants), every 4000th unit of
    the 4.5M -> 5.0M band; partitioned against gathers"""
    _, Bs, us = fc5.emit_units(4500001, 5000000)
    units = fc5.units_array(Bs[::4000], us[::4000])
    ctx.table_build(119047, 48, on_gpu=True)
    hdr = ctx.table_info()
    assert -(-(hdr.m_cells // 3) // (32768 * 256)) > 256
    part = ctx.run(units, fc5.MODE_BATCHED)  # probe_mode -1: partitioned for this table
    st = ctx.part_stats()
    direct = _run(ctx, units, probe_mode=0)
    say(f"part/flagship table: {st[1]}, probes={part.result['probes']}")
    assert st[1]["chunks"] > 0
    _same(direct, part)


def test_buffers_over_the_hbm_limit_fall_back_to_gathers(ctx):
    """option part_hbm_limit: by table size (-1) the search probes with gathers
    when the buffers would not fit, with the same results; forced (1) it fails"""
    _, Bs, us = fc5.emit_units(3200001, 3210000)
    units = fc5.units_array(Bs[::997], us[::997])
    ctx.table_build(76428, 48, on_gpu=True)
    ctx.set_option("part_hbm_limit", 1 << 30)
    try:
        out = ctx.run(units, fc5.MODE_BATCHED)
        assert ctx.part_stats()[1]["pieces"] == 0  # probed with gathers
        ctx.set_option("probe_mode", 1)
        with pytest.raises(fc5.Fc5Error):
            ctx.run(units, fc5.MODE_BATCHED)
    finally:
        ctx.set_option("probe_mode", -1)
        ctx.set_option("part_hbm_limit", 0)
    part = ctx.run(units, fc5.MODE_BATCHED)
    assert ctx.part_stats()[1]["pieces"] > 0
    _same(out, part)
