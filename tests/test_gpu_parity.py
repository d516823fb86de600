"""GPU parity: the HIP path (libfc5.so, both kernel forms) against the oracle.

Bit-exact bar: per-gate counters (calls, neg, zero, gated, probes), the filter
maybe-hit list (including false positives) and the final hit list must be
identical to the oracle on the same inputs and the same table bytes; planted
keys must all be recovered.  Tables are built by the product's host builder,
whose bytes are checked against the oracle's (themselves pinned to the
reference builder in test_oracle.py).
"""
import os
import sys

import numpy as np
import pytest

import fc5
import oracle as O

pytestmark = pytest.mark.gpu

COUNTERS = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = fc5.Context(0)
    yield c
    c.close()


def _units(lo, hi, every=1, offset=0):
    _, Bs, us = fc5.emit_units(lo, hi)
    return Bs[offset::every], us[offset::every]


def _oracle_run(Bs, us, table, side=None, use_gate=True):
    arr, ok = O.pack_units(Bs, us)
    assert ok.all()
    c, hits, nh = O.peel(arr, len(Bs), table, side=side, use_gate=use_gate, threads=min(16, os.cpu_count() or 8))
    return c, hits


def _load(ctx, N, r, extra=None):
    """same table bytes in product and oracle"""
    ctx.table_build(N, r, extra=extra)
    t = O.Table.build_pairs(N, r, extra=extra)
    h = ctx.table_info()
    assert (h.n_keys, h.m_cells, h.mix_seed) == (t.x.hdr.n_keys, t.x.hdr.m_cells, t.x.hdr.mix_seed)
    return t


def _compare(out, oc, ohits):
    for k in COUNTERS:
        assert out.result[k] == oc[k], (k, out.result[k], oc[k])
    got = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(got) == len(ohits)
    for f in ("unit", "d", "e", "flags", "fp"):
        assert np.array_equal(got[f], ohits[f]), f


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN, fc5.MODE_CHUNKED],
                         ids=["batched", "plain", "chunked"])
@pytest.mark.parametrize("r", [48, 16])
def test_band_100001_102000_full_matches_oracle(ctx, mode, r):
    say(f"B in [100001,102000] full r={r} mode={mode}")
    Bs, us = _units(100001, 102000)
    t = _load(ctx, 2428, r)
    out = ctx.run(fc5.units_array(Bs, us), mode)
    oc, oh = _oracle_run(Bs, us, t)
    _compare(out, oc, oh)
    assert out.result["calls"] == 10973276
    if r == 16:
        assert len(oh) > 0  # a non-empty false-positive list is exercised
    assert out.solutions == []


def test_band_100001_102000_no_gate(ctx):
    say("B in [100001,102000] no-gate")
    Bs, us = _units(100001, 100600)
    t = _load(ctx, 2428, 16)
    ctx.set_option("use_gate", 0)
    try:
        out = ctx.run(fc5.units_array(Bs, us))
    finally:
        ctx.set_option("use_gate", 1)
    oc, oh = _oracle_run(Bs, us, t, use_gate=False)
    _compare(out, oc, oh)
    assert out.result["gated"] == 0


@pytest.mark.parametrize("sweep", [0, 1], ids=["bitmap", "sieve"])
def test_sweep_forms_match_oracle(ctx, sweep):
    """both class-5 sweeps of the batched kernel (one LDS bitmap test per grid
    point; the bit sieve over d, 32 points a word) against the oracle: the full
    [100001,102000] band at r=16, sampled 3.2M units (B^6 > 2^128), without the
    gate, and with the strict gate"""
    say(f"sweep form {sweep}")
    t = _load(ctx, 2428, 16)
    ctx.set_option("sweep", sweep)
    try:
        for lo, hi, every in ((100001, 102000, 1), (3200001, 3210000, 499)):
            Bs, us = _units(lo, hi, every=every)
            out = ctx.run(fc5.units_array(Bs, us))
            oc, oh = _oracle_run(Bs, us, t)
            _compare(out, oc, oh)
            assert len(oh) > 0
        Bs, us = _units(100001, 100600)
        ctx.set_option("use_gate", 0)
        try:
            out = ctx.run(fc5.units_array(Bs, us))
        finally:
            ctx.set_option("use_gate", 1)
        oc, oh = _oracle_run(Bs, us, t, use_gate=False)
        _compare(out, oc, oh)
        Bs, us = _units(3200001, 3210000, every=499)
        arr, ok = O.pack_units(Bs, us)
        oc, oh, _ = O.peel(arr, len(Bs), t, strict=True, threads=min(16, os.cpu_count() or 8))
        ctx.set_option("strict_gate", 1)
        try:
            out = ctx.run(fc5.units_array(Bs, us))
        finally:
            ctx.set_option("strict_gate", 0)
        _compare(out, oc, oh)
        assert out.result["strict_out"] == oc["probes"] - oc["queries"]
    finally:
        ctx.set_option("sweep", 1)


@pytest.mark.parametrize("lo,hi,every", [(1000001, 1020000, 97), (2000001, 2010000, 211),
                                         (3200001, 3210000, 997)])
def test_sampled_units_match_oracle(ctx, lo, hi, every):
    """unit subsets of the 1.0M, 2.0M and 3.2M bands (B^6 > 2^128 in the last)"""
    say(f"sampled units [{lo},{hi}] every {every}")
    Bs, us = _units(lo, hi, every=every)
    t = _load(ctx, 2428, 16)
    b = ctx.band(fc5.units_array(Bs, us))
    outb = b.run(fc5.MODE_BATCHED)
    outp = b.run(fc5.MODE_PLAIN)
    b.close()
    oc, oh = _oracle_run(Bs, us, t)
    _compare(outb, oc, oh)
    _compare(outp, oc, oh)


def test_plants_in_main_table_recovered(ctx):
    say("plants in main table, B in [100001,102000]")
    Bs, us = _units(100001, 102000)
    pts = fc5.plant_points(Bs, us, 64, 0x615 << 8 | 2)
    keys = np.array([k for *_, k in pts], np.uint64)
    t = _load(ctx, 2428, 48, extra=keys)
    out = ctx.run(fc5.units_array(Bs, us))
    oc, oh = _oracle_run(Bs, us, t)
    _compare(out, oc, oh)
    found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits if h["flags"] & 1}
    assert {(i, e, d) for i, e, d, _ in pts} <= found
    assert out.solutions == []  # plants are not pair sums: the exact check rejects them


def test_plants_in_side_table_recovered(ctx):
    say("plants in side table, 3.2M sample")
    Bs, us = _units(3200001, 3210000, every=499)
    pts = fc5.plant_points(Bs, us, 32, 0x615 << 8 | 4)
    keys = np.array([k for *_, k in pts], np.uint64)
    _load(ctx, 2428, 48)
    ctx.side_build(keys)
    try:
        out = ctx.run(fc5.units_array(Bs, us))
    finally:
        ctx.side_clear()
    found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits if h["flags"] & 2}
    assert {(i, e, d) for i, e, d, _ in pts} <= found
    assert out.result["side_maybe"] >= 32


def test_device_query_matches_oracle(ctx):
    say("device query")
    t = _load(ctx, 2000, 16)
    rnd = np.random.default_rng(5)
    keys = rnd.integers(0, 2 ** 63, 50000, dtype=np.uint64) * 2 + 1
    pw = np.array([(x ** 6) % 2 ** 64 for x in range(1, 2001)], np.uint64)
    keys = np.concatenate([keys, pw[:500] + pw[1:501]])
    got = ctx.table_query(keys)
    want = np.array([t.contains(int(k)) for k in keys], np.uint8)
    assert np.array_equal(got, want)
    assert got[-500:].all()


def test_hit_overflow_is_reported(ctx):
    say("hit overflow")
    Bs, us = _units(100001, 100300)
    t = _load(ctx, 300, 8)
    ctx.set_option("hit_cap", 16)
    try:
        out = ctx.run(fc5.units_array(Bs, us), allow_overflow=True)
    finally:
        ctx.set_option("hit_cap", 1 << 20)
    oc, _ = _oracle_run(Bs, us, t)
    assert out.rc == fc5.E_OVERFLOW
    assert out.result["n_hits"] == oc["maybe"] > 16
    assert out.result["hits_dropped"] == oc["maybe"] - 16
    assert len(out.hits) == 16


def test_table_load_streamed_equals_whole_file(ctx, tmp_path):
    """fc5_table_load of a table above "stream_load_min" streams the file to the
    device in 64 MiB chunks (forced here on a 369 MB table: six chunks) and smokes
    it there: the device bytes equal the file's, the search gives the counters and
    hits of the table it was saved from, an FPR refusal and a wrong N keep the
    current table"""
    say("streamed table load")
    N = 10000
    ctx.table_build(N, 48, on_gpu=True)
    p = str(tmp_path / "t.bin")
    ctx.table_save(p)
    _, Bs, us = fc5.emit_units(400001, 420000)
    units = fc5.units_array(Bs[::5], us[::5])
    before = ctx.run(units)
    raw = open(p, "rb").read()
    q = str(tmp_path / "r4.bin")
    fc5.build_pairs_file(300, 4, q)  # far too many false positives: refused
    ctx.set_option("stream_load_min", 0)
    try:
        ctx.table_load(p, N)
        h, packed = ctx.table_download()
        assert len(raw) - 64 == len(packed) > 5 * (64 << 20)
        assert packed.tobytes() == raw[64:] and bytes(h) == raw[:64]
        after = ctx.run(units)
        for k in COUNTERS:
            assert after.result[k] == before.result[k], k
        assert np.array_equal(after.hits, before.hits)
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.table_load(q, 300)
        assert e.value.code == fc5.E_FPR
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.table_load(p, N + 1)
        assert e.value.code == fc5.E_NMISMATCH
        assert bytes(ctx.table_info()) == raw[:64]  # the loaded table stayed
        again = ctx.run(units)
        assert np.array_equal(again.hits, before.hits) and again.result["probes"] == before.result["probes"]
    finally:
        ctx.set_option("stream_load_min", 4 << 30)


def test_table_load_errors(ctx, tmp_path):
    say("table load errors")
    p = str(tmp_path / "t.bin")
    fc5.build_pairs_file(300, 48, p)
    ctx.table_load(p, 300)
    assert ctx.fpr_smoke() <= 16
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.table_load(p, 301)
    assert e.value.code == fc5.E_NMISMATCH
    q = str(tmp_path / "r8.bin")
    fc5.build_pairs_file(300, 8, q)
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.table_load(q, 300)
    assert e.value.code == fc5.E_FPR
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.table_load(str(tmp_path / "missing.bin"), 300)
    assert e.value.code == fc5.E_IO
    # damaged headers are refused before any allocation or copy
    raw = open(p, "rb").read()
    damaged = str(tmp_path / "damaged.bin")
    open(damaged, "wb").write(raw[:32] + (1 << 61).to_bytes(8, "little") + raw[40:])
    with pytest.raises(fc5.Fc5Error) as e:
        ctx.table_load(damaged, 300)
    assert e.value.code == fc5.E_IO
    hdr = fc5.StoreHeader.from_buffer_copy(raw[:64])
    for m in (1 << 61, 1 << 41):
        hdr.m_cells = m
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.table_upload(hdr, raw[64:])
        assert e.value.code == fc5.E_FORMAT, m
    ctx.table_load(p, 300)  # the context still works


def test_band_1000001_1020000_full_calls_and_64_plants(ctx):
    """the full 1.0M band (4.99e4 units): closed-form calls, counters conserved,
    every planted key recovered, plain and batched agree"""
    say("B in [1000001,1020000] full band")
    Bs, us = _units(1000001, 1020000)
    units = fc5.units_array(Bs, us)
    pts = fc5.plant_points(Bs, us, 64, 0x615 << 8 | 2)
    keys = np.array([k for *_, k in pts], np.uint64)
    ctx.table_build(2428, 48, extra=keys)
    b = ctx.band(units)
    out = b.run(fc5.MODE_BATCHED)
    say("batched", out.result)
    r = out.result
    assert r["calls"] == fc5.units_calls(units)
    assert r["calls"] == r["neg"] + r["zero"] + r["gated"] + r["probes"]
    found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits if h["flags"] & 1}
    assert {(i, e, d) for i, e, d, _ in pts} <= found
    outp = b.run(fc5.MODE_PLAIN)
    say("plain", outp.result)
    b.close()
    for k in COUNTERS:
        assert outp.result[k] == r[k], k
    assert np.array_equal(np.sort(outp.hits, order=["unit", "e", "d"]), np.sort(out.hits, order=["unit", "e", "d"]))


def _pair_keys(N):
    x = np.arange(1, N + 1, dtype=np.uint64)
    p = x * x * x
    p = p * p
    i, j = np.triu_indices(N)
    return p[i] + p[j]


@pytest.mark.parametrize("N,r", [(5000, 48), (3000, 16)])
def test_device_builder_valid_filter(ctx, N, r, tmp_path):
    """device-built table: same unique key count and geometry as the host
    builder (5000 > 4096 includes the systematic duplicates a^6 = 0 mod 2^64),
    no false negatives, FPR near 2^-r, and a save/reload round trip"""
    say(f"device builder N={N} r={r}")
    ctx.table_build(N, r, on_gpu=True)
    h = ctx.table_info()
    t = O.Table.build_pairs(N, r)
    assert (h.n_keys, h.m_cells, h.N, h.r) == (t.x.hdr.n_keys, t.x.hdr.m_cells, N, r)
    keys = _pair_keys(N)
    assert len(np.unique(keys)) == h.n_keys
    assert ctx.table_query(keys).all()
    fpr = ctx.fpr_smoke()
    assert fpr <= (16 if r == 48 else 40)
    p = str(tmp_path / "dev.bin")
    ctx.table_save(p)
    back = O.Table.load(p, N)
    assert back.header()["n_keys"] == h.n_keys
    sample = keys[:: max(1, len(keys) // 20000)]
    assert all(back.contains(int(k)) for k in sample)
    hh, packed = ctx.table_download()
    assert bytes(packed) == back.packed()


@pytest.mark.parametrize("r", [12, 24, 40, 64])
def test_other_cell_widths_match_oracle(ctx, r):
    """cells of 12, 24, 40 bits (straddling byte and word boundaries) and 64 bits:
    the uploaded host-built bytes are the oracle's and the search over them is
    bit-exact; a device-built table of the same width has every key and its
    search matches the oracle reading its downloaded bytes"""
    say(f"cell width r={r}")
    Bs, us = _units(100001, 101000)
    t = _load(ctx, 2428, r)
    h, packed = ctx.table_download()
    assert bytes(packed) == t.packed()
    out = ctx.run(fc5.units_array(Bs, us))
    oc, oh = _oracle_run(Bs, us, t)
    _compare(out, oc, oh)
    if r == 12:
        assert len(oh) > 0
    ctx.table_build(2428, r, on_gpu=True)
    assert ctx.table_query(_pair_keys(2428)).all()
    h, packed = ctx.table_download()
    td = O.Table.from_bytes({f: getattr(h, f) for f, _ in fc5.StoreHeader._fields_}, packed)
    out = ctx.run(fc5.units_array(Bs, us))
    oc, oh = _oracle_run(Bs, us, td)
    _compare(out, oc, oh)


def test_device_built_table_search_matches_oracle(ctx):
    """same bytes in both: the search over a device-built table, downloaded
    into the oracle, agrees bit-exactly (FP list included, r=16)"""
    say("search over device-built table")
    Bs, us = _units(100001, 102000)
    ctx.table_build(2428, 16, on_gpu=True)
    h, packed = ctx.table_download()
    t = O.Table.from_bytes({f: getattr(h, f) for f, _ in fc5.StoreHeader._fields_}, packed)
    out = ctx.run(fc5.units_array(Bs, us))
    oc, oh = _oracle_run(Bs, us, t)
    _compare(out, oc, oh)
    assert len(oh) > 0


F3_KEYS = ("jobs", "fail_t", "calls", "gated", "probes", "maybe")
C5_KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe")


def _oracle_mixed(cls, Bs, us, table):
    """per-class oracle counters and the hit list with caller unit indices"""
    out, hits = {}, []
    idx5 = np.nonzero(cls == 5)[0]
    if len(idx5):
        c5, h5 = _oracle_run(Bs[idx5], us[idx5], table)
        out[5] = c5
        h5 = h5.copy()
        h5["unit"] = idx5[h5["unit"]]
        hits.append(h5)
    idx3 = np.nonzero(cls != 5)[0]
    if len(idx3):
        for k in (2, 3, 4):
            sel = idx3[cls[idx3] == k]
            if not len(sel):
                continue
            c3, h3, _ = O.find3(cls[sel], Bs[sel], us[sel], table, threads=min(16, os.cpu_count() or 8))
            out[k] = {"jobs": c3["jobs"], "fail_t": c3["failT"], "calls": c3["calls"], "gated": c3["gated"],
                      "probes": c3["probes"], "maybe": c3["maybe"]}
            h3 = h3.copy()
            h3["unit"] = sel[h3["unit"]]
            hits.append(h3)
    allh = np.concatenate(hits) if hits else np.zeros(0, O.HIT_DTYPE)
    return out, np.sort(allh, order=["unit", "e", "d", "flags"])


def _compare_mixed(out, oc, oh):
    for k, ref in oc.items():
        got = out.by_class[k]
        for key in (C5_KEYS if k == 5 else F3_KEYS):
            assert got[key] == ref[key], (k, key, got[key], ref[key])
    got = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(got) == len(oh)
    for f in ("unit", "d", "e", "fp"):
        assert np.array_equal(got[f], oh[f]), f


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN, fc5.MODE_CHUNKED],
                         ids=["batched", "plain", "chunked"])
def test_mixed_classes_band_100001_102000_match_oracle(ctx, mode):
    """classes 2-5 of B in [100001,102000] in one band: cls2-4 run the find3
    leaf in the same persistent kernel as cls5 (batched) or the reference
    per-class launches (plain)"""
    say(f"mixed classes B in [100001,102000] mode={mode}")
    cls, Bs, us = fc5.emit_units(100001, 102000, classes=(2, 3, 4, 5))
    t = _load(ctx, 2428, 16)
    out = ctx.run(fc5.units_array(Bs, us), mode, cls=cls)
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare_mixed(out, oc, oh)
    assert sum(out.by_class[k]["calls"] for k in (2, 3, 4)) > 0 and len(oh) > 0


def test_mixed_classes_band_2000001_2010000_sample_match_oracle(ctx):
    """a unit sample of the 2.0M band, all four classes, batched and plain"""
    say("mixed classes B in [2000001,2010000] sample")
    cls, Bs, us = fc5.emit_units(2000001, 2010000, classes=(2, 3, 4, 5))
    sel = np.concatenate([np.nonzero(cls == k)[0][::m] for k, m in ((2, 60), (3, 60), (4, 400), (5, 400))])
    sel.sort()
    cls, Bs, us = cls[sel], Bs[sel], us[sel]
    t = _load(ctx, 2428, 16)
    b = ctx.band(fc5.units_array(Bs, us), cls)
    ob = b.run(fc5.MODE_BATCHED)
    op = b.run(fc5.MODE_PLAIN)
    b.close()
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare_mixed(ob, oc, oh)
    _compare_mixed(op, oc, oh)


def test_find3_plants_recovered(ctx):
    """planted three-core keys V = T(d) - c3^6 in the side table are found at (d, c3)"""
    say("find3 plants")
    cls, Bs, us = fc5.emit_units(2000001, 2000400, classes=(2, 3, 4))
    rng = np.random.default_rng(11)
    pts = []
    for i in rng.choice(len(Bs), 12, replace=False):
        k, B, u = int(cls[i]), int(Bs[i]), int(us[i])
        res, mod, fmax = O.free_d_cls234(B, u, k)
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
        pts.append((int(i), d, c3, (T - c3 ** 6) % 2 ** 64))
    _load(ctx, 2428, 48)
    ctx.set_option("use_gate", 0)  # plants need not pass the residue gate here
    ctx.side_build(np.array([p[3] for p in pts], np.uint64))
    try:
        out = ctx.run(fc5.units_array(Bs, us), cls=cls)
    finally:
        ctx.side_clear()
        ctx.set_option("use_gate", 1)
    found = {(int(h["unit"]), int(h["d"]), int(h["e"])) for h in out.hits if h["flags"] & 2}
    assert {(i, d, c) for i, d, c, _ in pts} <= found


def _same_run(x, y):
    for k in (2, 3, 4, 5):
        for key in F3_KEYS + ("neg", "zero", "side_maybe"):
            assert x.by_class[k][key] == y.by_class[k][key], (k, key)
    hx = np.sort(x.hits, order=["unit", "e", "d", "flags"])
    hy = np.sort(y.hits, order=["unit", "e", "d", "flags"])
    assert np.array_equal(hx, hy)


def test_graph_replays_match_eager(ctx):
    """the batched pipeline captured as a hipGraph, replayed several times with
    eager work between the replays, equals the eager result every time; an
    option change forces a re-capture"""
    say("graph replay")
    cls, Bs, us = fc5.emit_units(100001, 102000, classes=(2, 3, 4, 5))
    units = fc5.units_array(Bs, us)
    _load(ctx, 2428, 16)
    b = ctx.band(units, cls)
    other = ctx.band(units[::3], cls[::3])
    eager = b.run()
    eager_other = other.run()
    ctx.set_option("use_graph", 1)
    try:
        for rep in range(4):
            _same_run(b.run(), eager)
            ctx.set_option("use_graph", 0)
            _same_run(other.run(), eager_other)          # eager work between replays
            assert ctx.table_query(np.arange(1, 1000, dtype=np.uint64)).sum() >= 0
            ctx.set_option("use_graph", 1)
        ctx.set_option("use_gate", 0)
        ctx.set_option("use_graph", 0)
        eager_ng = b.run()
        ctx.set_option("use_graph", 1)
        _same_run(b.run(), eager_ng)                      # re-captured for the new option
        _same_run(b.run(), eager_ng)
    finally:
        ctx.set_option("use_gate", 1)
        ctx.set_option("use_graph", 0)
        b.close()
        other.close()


def test_tile_shapes_and_autotune_keep_results(ctx):
    """every rows-per-item instantiation, the choice by band size (0) and the
    tuned choice give the same counters and hits"""
    say("tile shapes / autotune")
    cls, Bs, us = fc5.emit_units(2000001, 2000300, classes=(2, 3, 4, 5))
    _load(ctx, 2428, 16)
    b = ctx.band(fc5.units_array(Bs, us), cls)
    base = b.run()
    try:
        for E in (4, 8, 16, 32, 0):
            ctx.set_option("rows_per_item", E)
            _same_run(b.run(), base)
        tune = b.autotune(sample_every=3)
        assert tune["rows_per_item"] in (4, 8, 16, 32) and tune["n_tried"] == 8
        assert tune["best_ms"] <= tune["default_ms"]
        _same_run(b.run(), base)
    finally:
        ctx.set_option("rows_per_item", 0)
        ctx.set_option("blocks_per_cu", 0)
        b.close()


def test_device_pair_search_exact(ctx):
    """fc5_pair_search against Python big-integer brute force: real pair sums
    (a < b, a = b, 1^6 + b^6), near misses, bounds that cut the larger term,
    zero; every reported pair is exact, also for targets near 2^127"""
    import random
    say("device pair search")
    rnd = random.Random(6162)
    targets, his = [], []
    for bits in (8, 12, 16):
        for _ in range(60):
            a, b = sorted(rnd.randrange(1, 2 ** bits) for _ in range(2))
            T = a ** 6 + b ** 6
            for t, h in ((T, b), (T, b - 1), (T + 1, b), (T - 1, b), (2 * b ** 6, b), (1 + b ** 6, b)):
                targets.append(t)
                his.append(max(h, 0))
    targets += [0, 2, 3]
    his += [10, 1, 5]
    got = ctx.pair_search(targets, his)
    want = [(i, a, b) for i, (T, h) in enumerate(zip(targets, his)) for a, b in brute_pairs(T, h)]
    assert got == want
    assert len(want) >= 500
    big = [2 ** 127 - 1, 2 * (2 ** 21 - 1) ** 6, (2 ** 21 - 5) ** 6 + 77 ** 6]
    got = ctx.pair_search(big, [2 ** 21] * 3)
    for i, a, b in got:
        assert a <= b and a ** 6 + b ** 6 == big[i]
    assert (1, 2 ** 21 - 1, 2 ** 21 - 1) in got and (2, 77, 2 ** 21 - 5) in got
    with pytest.raises(fc5.Fc5Error) as e:  # outside the exact range of the device sixth root
        ctx.pair_search([2 ** 127], [2 ** 21])
    assert e.value.code == fc5.E_ARG


def brute_pairs(T, hi):
    """all a <= b <= hi with a^6 + b^6 = T, by the larger term b"""
    import bigint_ref
    out = []
    b = min(hi, bigint_ref.iroot6(T))
    while b >= 1 and 2 * b ** 6 >= T:
        r = T - b ** 6
        a = bigint_ref.iroot6(r)
        if 1 <= a <= b and a ** 6 == r:
            out.append((a, b))
        b -= 1
    return sorted(out)


def test_cli_cls5_units_bbj_buc_and_exact_table(tmp_path):
    """the cls5 CLI reads Stage-1 units from a .bbj file as from a .buc file, and
    gives the same output when it checks survivors against the exact sorted table"""
    import re
    import subprocess
    from conftest import BIN
    say("cls5 CLI .bbj vs .buc, --exact-table")
    outs = []
    for ext, flags in (("buc", []), ("bbj", []), ("buc", ["--exact-table"])):
        p = str(tmp_path / f"u.{ext}")
        subprocess.run([os.path.join(BIN, "fourcore_hunt_v3"), "--lo", "100001", "--hi", "101000", "--classes", "5",
                        "--emit-units", p], check=True, capture_output=True)
        r = subprocess.run([os.path.join(BIN, "fourcore_cls5_gpu_v4"), "--units", p, "--N", "2428", "--r", "16",
                            "--quiet", "--counters"] + flags, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr[-1000:]
        lines = [l for l in r.stdout.splitlines() if l.startswith("COUNTERS") or l.startswith("----")]
        outs.append([re.sub(r"(kernel_ms|time)=[0-9.]+s?", "", l) for l in lines])
        if flags:
            assert "exact sorted pair table: 2948806 entries" in r.stderr
    assert outs[0] == outs[1] == outs[2] and outs[0]


def test_cli_cls5_strict_gate(tmp_path):
    """the cls5 CLI with --strict-gate: its counters are the oracle's strict peel
    on the same units and table bytes (r=16, the lowest width the CLI's FPR smoke
    accepts)"""
    import re
    import subprocess
    from conftest import BIN
    say("cls5 CLI --strict-gate")
    p = str(tmp_path / "u.buc")
    subprocess.run([os.path.join(BIN, "fourcore_hunt_v3"), "--lo", "100001", "--hi", "102000", "--classes", "5",
                    "--emit-units", p], check=True, capture_output=True)
    r = subprocess.run([os.path.join(BIN, "fourcore_cls5_gpu_v4"), "--units", p, "--N", "2428", "--r", "16",
                        "--quiet", "--counters", "--strict-gate"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-1000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("COUNTERS")][0]
    got = {k: int(float(v)) for k, v in re.findall(r"(\w+)=([0-9.]+)", line)}
    Bs, us = _units(100001, 102000)
    arr, ok = O.pack_units(Bs, us)
    oc, oh, _ = O.peel(arr, len(Bs), O.Table.build_pairs(2428, 16), strict=True,
                       threads=min(16, os.cpu_count() or 8))
    for k in ("calls", "neg", "zero", "gated", "probes", "maybe"):
        assert got[k] == oc[k], k
    assert got["strict_out"] == oc["probes"] - oc["queries"] > 0 and got["hits"] == len(oh)


def test_cli_xor_build_save_on_device(ctx, tmp_path):
    """xor_build_save --gpu writes a v2 .bin with the host builder's geometry and
    every key; the search over it (cls5 CLI, --load-table) matches the oracle
    reading the same file, FP maybe-hits included"""
    import re
    import subprocess
    from conftest import BIN
    say("xor_build_save --gpu")
    p = str(tmp_path / "dev.bin")
    r = subprocess.run([os.path.join(BIN, "xor_build_save"), "--gpu", "--N", "2428", "--r", "16", "--out", p],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-1000:]
    assert "(device)" in r.stderr and "[xor_build_save] wrote" in r.stdout
    t = O.Table.load(p, 2428)
    host = O.Table.build_pairs(2428, 16)
    hd = t.header()
    assert (hd["n_keys"], hd["m_cells"], hd["r"], hd["N"]) == (host.x.hdr.n_keys, host.x.hdr.m_cells, 16, 2428)
    ctx.table_load(p, 2428)
    assert ctx.table_query(_pair_keys(2428)).all()
    u = str(tmp_path / "u.buc")
    subprocess.run([os.path.join(BIN, "fourcore_hunt_v3"), "--lo", "100001", "--hi", "102000", "--classes", "5",
                    "--emit-units", u], check=True, capture_output=True)
    r = subprocess.run([os.path.join(BIN, "fourcore_cls5_gpu_v4"), "--units", u, "--N", "2428", "--r", "16",
                        "--load-table", p, "--quiet", "--counters"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-1000:]
    cls, Bs, us = fc5.emit_units(100001, 102000, classes=(5,))
    oc, oh = _oracle_run(Bs, us, t)
    line = next(l for l in r.stdout.splitlines() if l.startswith("COUNTERS"))
    got = dict((k, int(v)) for k, v in re.findall(r"(\w+)=(\d+)\b", line))
    for k in COUNTERS:
        assert got[k] == oc[k], (k, got[k], oc[k])
    assert oc["maybe"] > 0


@pytest.mark.parametrize("lo,hi,classes,band", [
    (100001, 102000, (2, 3, 4, 5), (0.0, 1.0)),
    (1, 3000, (2, 3, 4, 5), (0.0, 1.0)),
    (1000001, 1004000, (2, 3, 4, 5), (0.0, 1.0)),
    (2000001, 2002000, (4, 5), (0.2, 0.9)),
    (3500001, 3503000, (2, 3, 4, 5), (0.5, 1.0)),
    (9999001, 10000000, (2, 3, 5), (0.0, 1.0)),
])
def test_stage1_on_device_matches_host(ctx, lo, hi, classes, band):
    """Stage-1 emission on the device: the host emitter's units in its order"""
    say(f"device Stage-1 [{lo},{hi}] classes {classes} band {band}")
    dc, dB, du = ctx.emit_units_device(lo, hi, classes, band)
    hc, hB, hu = fc5.emit_units(lo, hi, classes=classes, u_band=band)
    assert len(hc) > 0
    assert np.array_equal(dc, hc) and np.array_equal(dB, hB) and np.array_equal(du, hu)


def test_band_from_range_equals_band_from_units(ctx):
    say("band from a B range")
    ctx.table_build(24285, 16, on_gpu=True)
    cls, Bs, us = fc5.emit_units(1000001, 1002000, classes=(4, 5))
    a = ctx.run(fc5.units_array(Bs, us), fc5.MODE_BATCHED, cls=cls)
    b = fc5.Band.from_range(ctx, 1000001, 1002000, classes=(4, 5))
    out = b.run(fc5.MODE_BATCHED)
    b.close()
    for k in COUNTERS:
        assert out.result[k] == a.result[k], k
    assert np.array_equal(out.hits, a.hits) and out.result["units"] == a.result["units"]


def _ragged_units(seed):
    """.buc lines beyond the Stage-1 list, as the reference reads them
    (fourcore_cls5_gpu_v4.cu:155-184): real units; (B, u) with an even u that is
    no seed lift (the residual is not a multiple of 42^6, so these run on the
    per-point 192-bit path), before and after the 2^128 wall; and lines that the
    reference's packing refuses (u >= B, odd u, u = 0, ...)"""
    import random

    rnd = random.Random(seed)
    Bs, us = [], []
    for lo in (100001, 1000001, 3200001):
        _, B1, u1 = fc5.emit_units(lo, lo + 60)
        Bs += [int(b) for b in B1[:4]]
        us += [int(u) for u in u1[:4]]
        for b in rnd.sample(sorted({int(x) for x in B1}), 3):
            for _ in range(3):
                Bs.append(b)
                us.append(2 * rnd.randint(b // 4, (b - 1) // 2))
    Bs += [100003, 100003, 100003, 100003, 1000003, 1000003]
    us += [100003, 100005, 7, 0, 999, 1000002]
    order = list(range(len(Bs)))
    rnd.shuffle(order)
    return [Bs[i] for i in order], [us[i] for i in order]


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN], ids=["batched", "plain"])
def test_ragged_and_irregular_units_match_oracle(ctx, mode):
    say(f"ragged unit list (regular, irregular, unpackable), mode={mode}")
    Bs, us = _ragged_units(17)
    arr, ok = O.pack_units(Bs, us)
    assert 0 < ok.sum() < len(Bs)
    t = _load(ctx, 2428, 16)
    out = ctx.run(fc5.units_array(Bs, us), mode)
    assert out.result["units_skipped"] == int((~ok).sum())
    assert out.result["units_irregular"] > 0
    keep = np.flatnonzero(ok)
    sub = (O.Unit5 * len(keep))(*[arr[int(i)] for i in keep])
    oc, oh = O.peel(sub, len(keep), t, threads=min(16, os.cpu_count() or 8))[:2]
    oh = oh.copy()
    oh["unit"] = keep[oh["unit"]].astype(np.uint32)
    oh = np.sort(oh, order=["unit", "e", "d", "flags"])
    _compare(out, oc, oh)
    assert len(oh) > 0


def test_units_either_side_of_the_2_pow_21_e_bound_match_oracle(ctx):
    """the top of the batched 128-bit T-space path: e's free range reaches 2^21
    at B = 14 * 2^21 = 29360128, so these Stage-1 units at B = 29360119 and
    29360123 (fmax1 = 2^21 - 1) run in the batched kernel and those at
    B = 29360141 and 29360143 on the per-point 192-bit path.  Units chosen so
    that the pair-sum gate passes some points (at B = 29360125 the residual is
    3, 4 or 5 mod 7 for every point, and sixth-power pair sums are 0, 1 or 2
    mod 7).  About 1.5e9 grid points per unit: the oracle needs ~90 s"""
    say("units either side of B = 14 * 2^21 (oracle ~90 s)")
    pick = [(29360119, 106578), (29360119, 14717196), (29360123, 11148),
            (29360141, 65340), (29360143, 165162), (29360143, 14658612)]
    Bs = np.array([b for b, _ in pick], np.uint64)
    us = np.array([u for _, u in pick], np.uint64)
    _, eB, eu = fc5.emit_units(29360119, 29360143)
    assert set(pick) <= set(zip(eB.tolist(), eu.tolist()))
    t = _load(ctx, 2428, 16)
    b = ctx.band(fc5.units_array(Bs, us))
    outb = b.run(fc5.MODE_BATCHED)
    outp = b.run(fc5.MODE_PLAIN)
    b.close()
    assert outb.result["units_irregular"] == 3
    oc, oh = _oracle_run(Bs, us, t)
    say("oracle", oc)
    _compare(outb, oc, oh)
    _compare(outp, oc, oh)
    assert oc["probes"] > 0 and len(oh) > 0


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN], ids=["batched", "plain"])
def test_empty_band(ctx, mode):
    say(f"empty band, mode={mode}")
    _load(ctx, 300, 48)
    out = ctx.run(fc5.units_array([], []), mode)
    assert all(out.result[k] == 0 for k in COUNTERS + ("n_hits", "units", "units_skipped"))
    assert len(out.hits) == 0 and out.solutions == []


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN], ids=["batched", "plain"])
def test_mixed_ragged_lines_match_oracle(ctx, mode):
    """.buc lines of every class in one band: Stage-1 units of classes 2-5;
    class 2-4 lines with other u (most free terms give a residual that is not a
    multiple of 42^6 and count as failT, fourcore_find_v4.cu:179-185); lines with
    u >= B (every free term fails); classes 1 and 6 (skipped)"""
    import random

    say(f"mixed ragged lines mode={mode}")
    rnd = random.Random(23)
    cls, Bs, us = fc5.emit_units(100001, 100600, classes=(2, 3, 4, 5))
    cls, Bs, us = [int(x) for x in cls], [int(x) for x in Bs], [int(x) for x in us]
    n_stage1 = len(cls)
    for _ in range(30):
        i = rnd.randrange(n_stage1)
        cls.append(rnd.choice((2, 3, 4)))
        Bs.append(Bs[i])
        us.append(rnd.randrange(1, Bs[i]))
    for k in (2, 3, 4):
        cls += [k, k]
        Bs += [100003, 100003]
        us += [100003 + k, 100003]
    cls += [1, 6, 1]
    Bs += [100003, 100003, 1000003]
    us += [6, 6, 12]
    order = list(range(len(cls)))
    rnd.shuffle(order)
    cls = np.array([cls[i] for i in order], np.int64)
    Bs = np.array([Bs[i] for i in order], np.int64)
    us = np.array([us[i] for i in order], np.int64)
    t = _load(ctx, 2428, 16)
    out = ctx.run(fc5.units_array(Bs, us), mode, cls=cls)
    oc, oh = _oracle_mixed(cls, Bs, us, t)
    _compare_mixed(out, oc, oh)
    assert out.result["units_skipped"] >= 3
    assert out.result["units_irregular"] > 0
    assert sum(oc[k]["fail_t"] for k in (2, 3, 4)) > 0 and len(oh) > 0


def test_band_item_count_limit(ctx):
    """a band whose work items would overflow the launch's 32-bit item counter is
    refused when it is created, not run with wrapped item numbers"""
    say("work-item limit")
    _, Bs, us = fc5.emit_units(29000003, 29000003)  # a regular unit with 17052 e-rows
    n = 1100000
    units = fc5.units_array(np.repeat(Bs[:1], n), np.repeat(us[:1], n))
    ctx.table_build(300, 48)
    ctx.set_option("rows_per_item", 4)  # 4263 items per unit: 4.7e9 in all
    try:
        with pytest.raises(fc5.Fc5Error) as e:
            ctx.band(units)
        assert e.value.code == fc5.E_ARG and "work items" in str(e.value)
        ctx.set_option("rows_per_item", 8)  # 2132 items per unit: 2.3e9, accepted (not run)
        ctx.band(units).close()
    finally:
        ctx.set_option("rows_per_item", 0)
