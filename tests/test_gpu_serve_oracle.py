"""GPU: the resident serving loop at its own workload, against the oracle.

One device pair table for B up to 3.52M (N = 83,809, r = 16: 4.32e9 cells, so
cell indices pass 2^32 and every false-positive list is non-empty), six bands
of classes 4 and 5 submitted staggered and served by continuous batching, with
starts on both sides of the two-limb wall B^6 = 2^128 (B = 2,642,245) up to
3.5M.  Each served band's per-class counters (class 5: calls, neg, zero,
gated, probes, maybe; class 4: jobs, fail_t, calls, gated, probes, maybe) and
its sorted maybe-hit list must equal the oracle's (oracle.peel, oracle.find3)
on the downloaded table bytes, and its result alone.  The table's 64-bit cell
indexing is checked directly (pair keys whose third cell lies past 2^32), and
both kernel forms search a unit sample of the top band against the oracle.
"""
import os
import sys
import time

import numpy as np
import pytest

import fc5
import oracle as O

pytestmark = pytest.mark.gpu

N_TABLE = 83809  # 42 * 83809 + 41 = 3,520,019 >= every B served here
BANDS = [(1000001, 1000300), (1800001, 1800200), (2642201, 2642300), (2900001, 2900060), (3300001, 3300050),
         (3499951, 3500000)]
C5_KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe")
F3_KEYS = ("jobs", "fail_t", "calls", "gated", "probes", "maybe")
M64 = (1 << 64) - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _threads():
    return min(16, os.cpu_count() or 8)


def mix64(x):
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & M64
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & M64
    x ^= x >> 33
    return x


@pytest.fixture(scope="module")
def served():
    """the table, the six bands served staggered, their results alone, and the
    table bytes wrapped for the oracle"""
    c = fc5.Context(0)
    t0 = time.time()
    c.table_build(N_TABLE, 16, on_gpu=True)
    hdr = c.table_info()
    say(f"serve/oracle: device table N={hdr.N} r={hdr.r} cells={hdr.m_cells} built in {time.time() - t0:.1f}s")
    assert hdr.m_cells > 2 ** 32
    bands = []
    for lo, hi in BANDS:
        cls, Bs, us = fc5.emit_units(lo, hi, classes=(4, 5))
        bands.append((cls, Bs, us, fc5.units_array(Bs, us)))
    alone = [c.run(u, fc5.MODE_BATCHED, cls=cl) for cl, _, _, u in bands]
    c.serve_start(max_bands=3, round_points=2e9)
    try:
        ids = []
        for k, (cl, _, _, u) in enumerate(bands):
            ids.append(c.serve_submit(u, cl))
            if k % 2:
                time.sleep(0.02)
        for i in ids:
            assert c.serve_wait(i, 300.0)
        outs = [c.serve_result(i) for i in ids]
        timing = [c.serve_timing(i) for i in ids]
        stats = c.serve_stats()
    finally:
        c.serve_stop()
    t0 = time.time()
    h, packed = c.table_download()
    table = O.Table.from_bytes({f: getattr(h, f) for f, _ in fc5.StoreHeader._fields_}, packed)
    say(f"serve/oracle: 6 bands served in {stats['rounds']} rounds; table bytes downloaded in {time.time() - t0:.1f}s")
    yield {"ctx": c, "bands": bands, "alone": alone, "outs": outs, "timing": timing, "stats": stats,
           "table": table, "hdr": h}
    c.close()


def _oracle_band(cls, Bs, us, table):
    """per-class oracle counters and the sorted maybe-hit list (caller unit indices)"""
    out, hits = {}, []
    i5 = np.flatnonzero(cls == 5)
    arr, ok = O.pack_units(Bs[i5], us[i5])
    assert ok.all()
    c5, h5, _ = O.peel(arr, len(i5), table, threads=_threads())
    out[5] = c5
    h5 = h5.copy()
    h5["unit"] = i5[h5["unit"].astype(np.int64)]
    hits.append(h5)
    i4 = np.flatnonzero(cls == 4)
    c4, h4, _ = O.find3(cls[i4], Bs[i4], us[i4], table, threads=_threads())
    out[4] = {"jobs": c4["jobs"], "fail_t": c4["failT"], "calls": c4["calls"], "gated": c4["gated"],
              "probes": c4["probes"], "maybe": c4["maybe"]}
    h4 = h4.copy()
    h4["unit"] = i4[h4["unit"].astype(np.int64)]
    hits.append(h4)
    return out, np.sort(np.concatenate(hits), order=["unit", "e", "d", "flags"])


def _compare(out, oc, oh):
    for k, keys in ((5, C5_KEYS), (4, F3_KEYS)):
        for key in keys:
            assert out.by_class[k][key] == oc[k][key], (k, key, out.by_class[k][key], oc[k][key])
    got = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(got) == len(oh), (len(got), len(oh))
    for f in ("unit", "d", "e", "fp"):
        assert np.array_equal(got[f], oh[f]), f
    assert np.array_equal(got["flags"] & 3, oh["flags"] & 3)


def test_bands_overlapped(served):
    t = served["timing"]
    assert served["stats"]["done"] == len(BANDS)
    assert any(t[j]["t_admit"] < t[k]["t_done"] and t[j]["t_done"] > t[k]["t_admit"]
               for j in range(len(t)) for k in range(j))


@pytest.mark.parametrize("k", range(len(BANDS)), ids=[f"B{lo}-{hi}" for lo, hi in BANDS])
def test_served_band_matches_oracle(served, k):
    lo, hi = BANDS[k]
    cls, Bs, us, _ = served["bands"][k]
    out = served["outs"][k]
    say(f"serve/oracle: band [{lo},{hi}] {len(Bs)} units, {out.result['calls']} points+c3 values")
    t0 = time.time()
    oc, oh = _oracle_band(cls, Bs, us, served["table"])
    say(f"serve/oracle:   oracle {time.time() - t0:.1f}s, {len(oh)} maybe-hits")
    assert len(oh) > 0  # r = 16: the false-positive list is exercised
    _compare(out, oc, oh)
    alone = served["alone"][k]
    for c in (4, 5):
        for key in C5_KEYS + ("jobs", "fail_t", "side_maybe"):
            assert out.by_class[c][key] == alone.by_class[c][key], (c, key)
    assert np.array_equal(out.hits, alone.hits) and out.solutions == alone.solutions == []
    assert out.by_class[5]["calls"] == fc5.units_calls(served["bands"][k][3][cls == 5])


def test_cells_past_2_32_indexed_exactly(served):
    """pair keys whose third cell index is at least 2^32 are present on the device
    and in the oracle's view of the same bytes"""
    h = served["hdr"]
    block = h.m_cells // 3
    rng = np.random.default_rng(83809)
    keys = []
    while len(keys) < 400:
        a, b = sorted(int(x) for x in rng.integers(1, N_TABLE + 1, 2))
        key = (a ** 6 + b ** 6) & M64
        hh = mix64((key + h.mix_seed) & M64)
        if mix64(hh ^ 0xBF58476D1CE4E5B9) % block + 2 * block >= 2 ** 32:
            keys.append(key)
    keys = np.array(keys, np.uint64)
    assert served["ctx"].table_query(keys).all()
    assert all(served["table"].contains(int(k)) for k in keys)


@pytest.mark.parametrize("mode", [fc5.MODE_BATCHED, fc5.MODE_PLAIN], ids=["batched", "plain"])
def test_search_over_table_past_2_32_matches_oracle(served, mode):
    """a class-5 unit sample of the top band, both kernel forms, same table bytes"""
    cls, Bs, us, _ = served["bands"][-1]
    i5 = np.flatnonzero(cls == 5)[::8]
    out = served["ctx"].run(fc5.units_array(Bs[i5], us[i5]), mode)
    arr, ok = O.pack_units(Bs[i5], us[i5])
    oc, oh, _ = O.peel(arr, len(i5), served["table"], threads=_threads())
    for key in C5_KEYS:
        assert out.result[key] == oc[key], key
    got = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(got) == len(oh) > 0
    for f in ("unit", "d", "e", "fp"):
        assert np.array_equal(got[f], oh[f]), f
