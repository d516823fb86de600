#!/usr/bin/env python3
"""Benchmark of the class-5 free peel on one MI355X.

A step is one pass of the batched persistent peel over one band of Stage-1
class-5 units, with every input (units, pair table, planted side table)
resident in HBM.  Default workload (--config 4): eligible B in
[3200001, 3210000] (B^6 > 2^128), cls5 units of that band, the xor pair table
for N = 3210000 // 42 = 76428 at r = 48 (built on the device before timing),
256 planted keys in a side table, all of which must be recovered every step.

Printed: one JSON line with the metric ((d,e) grid points per second), the
eligible B cleared per hour, the plain one-launch-per-chunk form on a unit
subset, the roofline of the dominant kernel (k_cls5_batched: algorithmic
gather bytes per launch over its event-timed duration) and the CPU oracle
timed on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "six-one-five_amd", "python"))

import numpy as np  # noqa: E402

import fc5  # noqa: E402

METRIC = "cls5 (d,e) probes/sec + eligible B cleared/hour at B≈3.2M, 1 GPU, bit-exact"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    1: dict(lo=100001, hi=102000, N=2428, plants=0, plant_mode=None, seed=0x615 << 8 | 1, classes=(5,)),
    2: dict(lo=1000001, hi=1020000, N=24285, plants=64, plant_mode="main", seed=0x615 << 8 | 2, classes=(5,),
            exact=True),
    3: dict(lo=2000001, hi=2010000, N=47857, plants=0, plant_mode=None, seed=0x615 << 8 | 3,
            classes=(2, 3, 4, 5)),
    4: dict(lo=3200001, hi=3210000, N=76428, plants=256, plant_mode="side", seed=0x615 << 8 | 4, classes=(5,),
            exact=True, from_file=True),
}
COUNTER_KEYS = ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe", "jobs", "fail_t")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Model name and logical CPU count of this host (reported beside the CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def eligible_count(lo, hi):
    return sum(1 for B in range(lo, hi + 1) if B % 2 and B % 3 and B % 7)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=4, choices=sorted(CONFIGS) + [5, 617],
                    help="1-4: class-5 peel bands; 5: serving (Poisson band arrivals); 617: the (6,1,7) twin")
    ap.add_argument("--band", type=int, nargs=2, metavar=("B_LO", "B_HI"), default=None,
                    help="1-4: run the config's classes over this B range instead (table N = B_HI // 42, "
                         "no planted keys), e.g. a band whose reference run time is published")
    ap.add_argument("--table-n", type=int, default=0,
                    help="with --band: pair table N (default B_HI // 42), e.g. to run a long band as pieces "
                         "over the table of the whole band")
    ap.add_argument("--serve-requests", type=int, default=200, help="5: number of band requests")
    ap.add_argument("--serve-load", type=float, default=0.8, help="5: offered load as a fraction of capacity")
    ap.add_argument("--serve-calibrate", type=int, default=24, help="5: bands of the closed-loop capacity run")
    ap.add_argument("--serve-check", type=int, default=8, help="5: bands re-run alone and compared")
    ap.add_argument("--twin-lo", type=int, default=400471, help="617: first B")
    ap.add_argument("--twin-hi", type=int, default=400471, help="617: last B")
    ap.add_argument("--twin-cpu-every", type=int, default=128, help="617: CPU oracle on every k-th candidate")
    ap.add_argument("--twin-plain-every", type=int, default=4, help="617: plain form on every k-th candidate")
    ap.add_argument("--r", type=int, default=48)
    ap.add_argument("--exact-table", choices=["auto", "on", "off"], default="auto",
                    help="exact sorted pair table beside the xor filter, planted targets in both, hits "
                         "checked by lookup (auto: on for --config 2 and 4)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    ap.add_argument("--no-strict", action="store_true",
                    help="skip the strict-gate leg (the same band with the option strict_gate on)")
    ap.add_argument("--plain-every", type=int, default=64, help="plain form on every k-th unit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-every", type=int, default=0,
                    help="CPU oracle on every k-th unit (default: k for about 1e10 grid points, ~15 s on 16 threads)")
    ap.add_argument("--cpu-offset", type=int, default=0, help="first unit of the CPU oracle sample")
    ap.add_argument("--table-dir", default=None,
                    help="4: directory for the table file written and loaded back before timing (default: TMPDIR)")
    ap.add_argument("--no-table-file", action="store_true",
                    help="4: search the device-built table directly instead of one loaded from a file")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--no-measure", action="store_true", help="skip the gather/copy microbenchmarks")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shape")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    ap.add_argument("--probe", choices=["auto", "gathers", "partitioned"], default="auto",
                    help="filter probes: three gathers each, or partitioned by table slice (auto: by table)")
    ap.add_argument("--no-gathers", action="store_true",
                    help="skip the comparison leg with three gathers per probe (partitioned runs)")
    ap.add_argument("--load-policy", type=int, default=None, choices=[0, 1, 2, 3],
                    help="cache policy of the cell gathers (library default: by table size)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: a band shards over GPUs by units, each GPU holding its own "
              "copy of the read-only pair table, with only counters and hit records gathered at the end; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == 617:
        return bench_twin(args)
    if args.config == 5:
        return bench_serve(args)
    cfg = dict(CONFIGS[args.config])
    if args.band:
        lo, hi = args.band
        if not 1 <= lo <= hi:
            raise SystemExit("--band needs 1 <= B_LO <= B_HI")
        if args.table_n and args.table_n < hi // 42:
            raise SystemExit("--table-n must be at least B_HI // 42")
        cfg.update(lo=lo, hi=hi, N=args.table_n or hi // 42, plants=0, plant_mode=None, exact=False)
    t_setup = time.time()
    mixed = cfg["classes"] != (5,)
    ucls, Bs, us = fc5.emit_units(cfg["lo"], cfg["hi"], classes=cfg["classes"])
    units = fc5.units_array(Bs, us)
    n_B = eligible_count(cfg["lo"], cfg["hi"])
    calls_closed = fc5.units_calls(units[ucls == 5]) if mixed else fc5.units_calls(units)
    log(f"[bench] band: B=[{cfg['lo']},{cfg['hi']}] eligible_B={n_B} units={len(units)} "
        f"calls={calls_closed:.4e}")

    ctx = fc5.Context(0)
    if args.blocks_per_cu:
        ctx.set_option("blocks_per_cu", args.blocks_per_cu)
    if args.load_policy is not None:
        ctx.set_option("load_policy", args.load_policy)
    use_exact = args.exact_table == "on" or (args.exact_table == "auto" and cfg.get("exact", False))
    # plants that pass the strict gate too, as a real pair sum does: both legs must find them all
    plants = fc5.plant_points(Bs, us, cfg["plants"], cfg["seed"], exact=True, strict=True) if cfg["plants"] else []
    pkeys = np.array([p[3] for p in plants], np.uint64)
    t0 = time.time()
    ctx.table_build(cfg["N"], args.r, extra=pkeys if cfg["plant_mode"] == "main" else None, on_gpu=True)
    hdr = ctx.table_info()
    t_build = time.time() - t0
    log(f"[bench] device table N={hdr.N} keys={hdr.n_keys} cells={hdr.m_cells} "
        f"packed={fc5_packed_gb(hdr):.2f} GB built in {t_build:.1f}s")
    table_file = None
    if cfg.get("from_file") and not args.band and not args.no_table_file:
        table_file = table_via_file(ctx, hdr, args.table_dir)
        hdr = ctx.table_info()
    fpr = ctx.fpr_smoke()
    if fpr > 16:
        raise SystemExit(f"FPR smoke {fpr} > 16: refuse")
    # no false negatives: 2^20 random pair keys a^6 + b^6 (mod 2^64), a <= b <= N, are all present
    rng = np.random.default_rng(cfg["seed"] ^ 0xF00D)
    ab = np.sort(rng.integers(1, cfg["N"] + 1, size=(1 << 20, 2), dtype=np.uint64), axis=1)
    p6 = ab * ab * ab
    p6 = p6 * p6
    present = ctx.table_query(p6[:, 0] + p6[:, 1])
    assert present.all(), f"{int((~present.astype(bool)).sum())} pair keys missing from the table"
    log(f"[bench] table check: FPR smoke {fpr}/200000, 2^20 random pair keys all present")
    if cfg["plant_mode"] == "side":
        ctx.side_build(pkeys, args.r)
    exact = None
    if use_exact:
        t1 = time.time()
        ctx.exact_build(cfg["N"], [p[4] for p in plants] if cfg["plant_mode"] == "main" else [])
        n_ex, _, s_ex = ctx.exact_info()
        exact = {"entries": n_ex, "GB": n_ex * 16 / 1e9, "build_s": time.time() - t1,
                 "verify": "maybe-hits looked up in the exact table (equal range)"}
        # spot checks: random pair sums are found, and windows across the table are strictly increasing
        rng = np.random.default_rng(cfg["seed"])
        ab = [tuple(sorted(map(int, rng.integers(1, cfg["N"] + 1, 2)))) for _ in range(256)]
        got = ctx.exact_lookup([x ** 6 + y ** 6 for x, y in ab], [cfg["N"]] * len(ab))
        assert sorted(got) == sorted((i, x, y) for i, (x, y) in enumerate(ab)), "exact table lookup"
        wl = min(1024, n_ex)
        for off in np.linspace(0, n_ex - wl, 16).astype(np.int64):
            w = ctx.exact_read(int(off), wl)
            assert all(w[i] < w[i + 1] for i in range(wl - 1)), f"exact table order at {off}"
        exact["spot_checks"] = f"256 random pair sums found; 16 windows of {wl} entries strictly increasing"
        ctx.set_option("verify_table", 1)
        log(f"[bench] exact sorted pair table: {n_ex} entries, {exact['GB']:.2f} GB, built in {exact['build_s']:.1f}s")
    t1 = time.time()
    band = ctx.band(units, ucls if mixed else None)
    log(f"[bench] band packed and uploaded in {time.time() - t1:.1f}s")
    plant_set = {(p[0], p[1], p[2]) for p in plants}
    flag = 2 if cfg["plant_mode"] == "side" else 1

    def check(out):
        r = out.result
        c5 = out.by_class[5]["calls"]
        assert c5 == calls_closed, (c5, calls_closed)
        assert r["calls"] == r["neg"] + r["zero"] + r["gated"] + r["probes"]
        assert r["hits_dropped"] == 0
        found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits if h["flags"] & flag}
        missing = plant_set - found
        assert not missing, f"{len(missing)} planted keys not recovered"
        if use_exact and cfg["plant_mode"] == "main":  # and confirmed by the exact table
            exact_found = {(int(h["unit"]), int(h["e"]), int(h["d"])) for h in out.hits
                           if h["flags"] & fc5.HIT_EXTRA}
            assert exact_found == plant_set, "planted targets not confirmed by the exact table"
        return len(plant_set & found)

    tune = None
    if not args.no_autotune:
        t1 = time.time()
        # trials on every k-th unit: 1/16 of the band, or fewer units for bands far above 2e12 grid points
        every = max(16, int(calls_closed / 2e12))
        tune = band.autotune(sample_every=every)
        tune["seconds"] = time.time() - t1
        tune["sample_every"] = every
        log(f"[bench] autotune: rows_per_item={tune['rows_per_item']} blocks_per_cu={tune['blocks_per_cu']} "
            f"({tune['best_ms']:.1f} ms vs default {tune['default_ms']:.1f} ms on 1/{every} of the band)")
    if not args.no_graph:
        ctx.set_option("use_graph", 1)
    ctx.set_option("probe_mode", PROBE_MODES[args.probe])
    for _ in range(args.warmup):
        check(band.run(fc5.MODE_BATCHED))
    log(f"[bench] setup {time.time() - t_setup:.1f}s; timing {args.steps} steps")
    kms, wms, outs = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = band.run(fc5.MODE_BATCHED)
        outs.append(out)
        kms.append(out.result["kernel_ms"])
        wms.append(out.result["wall_ms"])
    wall = time.perf_counter() - t0
    recovered = [check(o) for o in outs]
    r = outs[-1].result
    for o in outs[1:]:
        for k in COUNTER_KEYS:
            assert o.result[k] == outs[0].result[k], k
    ms_step = 1000.0 * wall / args.steps
    value = r["calls"] / (ms_step / 1000.0)
    kern_s = float(np.mean(kms)) / 1000.0
    cell_bytes = math.ceil(args.r / 8)
    alg_bytes = r["probes"] * 3 * cell_bytes
    pst = ctx.part_stats()
    partitioned = pst[1]["pieces"] > 0
    log(f"[bench] {ms_step:.1f} ms/step (library {float(np.mean(wms)):.1f} ms), kernel {1000 * kern_s:.1f} ms, "
        f"{value:.4e} points/s, "
        f"probes={r['probes']} maybe={r['maybe']} side_maybe={r['side_maybe']}"
        + (f"; partitioned probing: emit {pst[0]['emit']:.0f} ms, pipeline {pst[0]['pipeline']:.0f} ms, "
           f"identify {pst[0]['identify']:.0f} ms, {pst[1]['pieces']} pieces" if partitioned else ""))
    part = partition_figures(pst, r, hdr, ms_step) if partitioned else None

    gathers = None
    if partitioned and not args.no_gathers:
        gathers = gathers_leg(ctx, band, outs[-1], check, ms_step, cell_bytes, cfg)

    strict = None
    if not args.no_strict:
        saved = (tune["rows_per_item"], tune["blocks_per_cu"]) if tune else (0, args.blocks_per_cu)
        strict = strict_leg(ctx, band, outs[-1], Bs, us, ucls if mixed else None, args, check, value, saved, n_B)

    ctx.set_option("use_graph", 0)
    ctx.set_option("verify_table", 0)
    plain = None
    if not args.no_plain:
        # three gathers per probe in all three forms: batched/chunked isolates
        # the persistent single launch of the same kernel
        ctx.set_option("probe_mode", 0)
        sub = units[:: args.plain_every]
        b2 = ctx.band(sub, ucls[:: args.plain_every] if mixed else None)
        ob = b2.run(fc5.MODE_BATCHED)
        t1 = time.perf_counter()
        op = b2.run(fc5.MODE_PLAIN)
        tp = time.perf_counter() - t1
        t1 = time.perf_counter()
        ob2 = b2.run(fc5.MODE_BATCHED)
        tb = time.perf_counter() - t1
        t1 = time.perf_counter()
        oc = b2.run(fc5.MODE_CHUNKED)
        tc = time.perf_counter() - t1
        b2.close()
        for cl in (2, 3, 4, 5):
            for k in COUNTER_KEYS:
                assert op.by_class[cl][k] == ob.by_class[cl][k] == ob2.by_class[cl][k] == oc.by_class[cl][k], (cl, k)
        hb = np.sort(ob.hits, order=["unit", "e", "d", "flags"])
        hp = np.sort(op.hits, order=["unit", "e", "d", "flags"])
        hc = np.sort(oc.hits, order=["unit", "e", "d", "flags"])
        assert np.array_equal(hb, hp) and np.array_equal(hb, hc), "plain, chunked and batched hit lists differ"
        pv, bv, cv = op.result["calls"] / tp, ob2.result["calls"] / tb, oc.result["calls"] / tc
        plain = {"subset": f"every {args.plain_every}th unit ({len(sub)} units)", "points_per_s": pv,
                 "batched_points_per_s_same_subset": bv,
                 "chunked_points_per_s_same_subset": cv,
                 "eligible_B_per_hour": n_B * 3600.0 * pv / r["calls"],
                 "speedup_batched_vs_plain": value / pv,
                 "speedup_batched_vs_plain_same_subset": bv / pv,
                 "speedup_batched_vs_chunked_same_subset": bv / cv,
                 "speedup_chunked_vs_plain_same_subset": cv / pv,
                 "chunked": "the T-space kernel launched once per 16 units with a synchronize after each "
                            "(the reference's launch pattern): batched/chunked isolates the persistent "
                            "single launch, chunked/plain the T-space arithmetic",
                 "counters_identical": True}
        ctx.set_option("probe_mode", -1)
        log(f"[bench] plain form {pv:.4e} points/s on the subset; batched/plain = {value / pv:.1f}x; "
            f"same subset: batched {bv:.4e}, chunked {cv:.4e} (batched/chunked {bv / cv:.2f}x, "
            f"chunked/plain {cv / pv:.1f}x)")

    ceiling = None
    if not args.no_measure:
        ceiling = measure_ceilings(ctx, hdr, r["probes"] / kern_s)

    traffic, valu = None, None
    if not args.band:
        tprof = load_profile(f"traffic_cls5_batched_B{cfg['lo']}-{cfg['hi']}.json")
        traffic = tprof.get("traffic_bytes_per_launch") if tprof else None
        valu = valu_roofline(load_profile(f"valu_cls5_batched_B{cfg['lo']}-{cfg['hi']}.json"), ceiling)
    # the gather kernel's roofline: the launch timed here (three gathers per
    # probe), or the gathers leg's launch when this run probed by partition
    g_s = (gathers["kernel_ms"] / 1000.0) if gathers else (None if partitioned else kern_s)
    gather_roof = None
    if g_s:
        gather_roof = {"bound": "hbm", "achieved": alg_bytes / g_s / 1e9, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": alg_bytes / g_s / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                       "kernel": "k_cls5_batched",
                       "algorithmic_bytes_per_launch": alg_bytes,
                       "per_unit": f"{3 * cell_bytes} B per filter probe (3 cells x {cell_bytes} B)",
                       "line_fill_GBps_at_128B": r["probes"] * 3 * 128 / g_s / 1e9,
                       # measured HBM traffic (128-byte lines per random gather) over the same launch time
                       "traffic_GBps": traffic / g_s / 1e9 if traffic else None,
                       "traffic_frac": traffic / g_s / 1e9 / HBM_PEAK_GBS if traffic else None,
                       "valu": valu}
    if partitioned:
        lt = None if args.band else load_profile(f"traffic_part_lookup_B{cfg['lo']}-{cfg['hi']}.json")
        part["roofline"]["traffic"] = lt.get("traffic_bytes_per_launch") if lt else None
        if part["roofline"]["traffic"]:
            lk_s = part["roofline"]["avg_launch_ms"] / 1000.0
            part["roofline"]["traffic_GBps"] = part["roofline"]["traffic"] / lk_s / 1e9
            part["roofline"]["traffic_frac"] = part["roofline"]["traffic_GBps"] / HBM_PEAK_GBS
        if gathers:
            gathers["roofline"] = gather_roof

    cpu = cpu_ref = None
    if not args.no_cpu_baseline:
        every = args.cpu_every or max(1, round(r["calls"] / 1e10))
        cpu = cpu_baseline(ctx, units, ucls, every, cfg, r, pkeys, args.cpu_offset)
        if not args.band:
            cpu_ref = cpu_reference(args.config)

    line = {
        "metric": METRIC, "value": value, "unit": "(d,e) grid points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (deterministic Stage-1 enumeration of the band"
                + ("; seeded planted keys)" if plants else ")"),
        "config": {"workload": "batched persistent cls5 peel"
                               + (" with partitioned filter probing" if partitioned else "") + ", eligible B in "
                               f"[{cfg['lo']},{cfg['hi']}] ({n_B} B, {len(units)} units, "
                               f"{r['calls']} grid points), xor pair table N={hdr.N} r={hdr.r} "
                               f"({fc5_packed_gb(hdr):.2f} GB packed, "
                               + ("device-built, written to a .bin file and loaded from it" if table_file and
                                  table_file.get("loaded") else "device-built") + "), "
                               f"{len(plants)} planted keys ({cfg['plant_mode'] or 'none'} table)",
                   "B_lo": cfg["lo"], "B_hi": cfg["hi"], "N": int(hdr.N), "r": int(hdr.r),
                   "units": int(len(units)), "parallelism": "single GPU"},
        "eligible_B_per_hour": n_B * 3600.0 / (ms_step / 1000.0),
        "reference_gpu_projection": reference_projection(cfg, n_B * 3600.0 / (ms_step / 1000.0)),
        "filter_probes_per_s": r["probes"] / (ms_step / 1000.0),
        "cell_gathers_per_s": 3 * r["probes"] / (ms_step / 1000.0),
        "counters": {k: r[k] for k in COUNTER_KEYS + ("n_hits", "solutions")},
        "counters_by_class": {str(cl): {k: outs[-1].by_class[cl][k] for k in COUNTER_KEYS}
                              for cl in cfg["classes"]} if mixed else None,
        "plants_recovered": f"{min(recovered)}/{len(plants)}",
        "kernel_ms": float(np.mean(kms)),
        "roofline": part["roofline"] if partitioned else gather_roof,
        # the filter's algorithmic bytes (3 cells of every probe) over the whole step
        "probe_roofline": {"achieved": alg_bytes / (ms_step / 1000.0) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": alg_bytes / (ms_step / 1000.0) / 1e9 / HBM_PEAK_GBS,
                           "per_unit": f"{3 * cell_bytes} B per filter probe (3 cells x {cell_bytes} B)",
                           "probes_per_s": r["probes"] / (ms_step / 1000.0)},
        "partitioned_probing": part,
        "gathers": gathers,
        "gather_ceiling": ceiling,
        "autotune": tune,
        "graph": not args.no_graph,
        "plain_form": plain,
        "strict_gate": strict,
        "cpu_baseline": cpu,
        "cpu_reference": cpu_ref,
        "table_build_s": t_build,
        "table_file": table_file,
        "exact_table": exact,
    }
    print(json.dumps(line), flush=True)
    band.close()
    ctx.close()


def bench_twin(args):
    """(6,1,7) class-5 twin: every Branch-A candidate of B in [twin_lo, twin_hi]
    in one persistent launch per step, table N = twin_hi // 42 at r = 48."""
    lo, hi = args.twin_lo, args.twin_hi
    Bs = [B for B in range(lo, hi + 1) if B % 2 and B % 3 and B % 7]
    cands = np.concatenate([fc5.twin_cands(B) for B in Bs])
    N = hi // 42
    ctx = fc5.Context(0)
    if args.load_policy is not None:
        ctx.set_option("load_policy", args.load_policy)
    if args.blocks_per_cu:
        ctx.set_option("blocks_per_cu", args.blocks_per_cu)
    t0 = time.time()
    ctx.table_build(N, args.r, on_gpu=True)
    hdr = ctx.table_info()
    log(f"[bench] 617 twin: B=[{lo},{hi}] eligible_B={len(Bs)} candidates={len(cands)}; device table N={N} "
        f"keys={hdr.n_keys} packed={fc5_packed_gb(hdr):.3f} GB built in {time.time() - t0:.1f}s")
    known = {(400471, (13860, 83202, 93240, 129696, 144756, 355502, 357609))}
    for _ in range(args.warmup):
        res, hits, sols = ctx.twin_run(cands)
    outs, kms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, hits, sols = ctx.twin_run(cands)
        outs.append(res)
        kms.append(res["kernel_ms"])
    wall = time.perf_counter() - t0
    for o in outs[1:]:
        for k in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "maybe"):
            assert o[k] == outs[0][k], k
    got = {(s["B"], tuple(s["terms"])) for s in sols if s["gcd"] == 1}
    want = {k for k in known if lo <= k[0] <= hi}
    assert want <= got, "known solution not recovered"
    ms_step = 1000.0 * wall / args.steps
    kern_s = float(np.mean(kms)) / 1000.0
    calls_s = res["calls"] / (ms_step / 1000.0)
    cell_bytes = math.ceil(args.r / 8)
    alg_bytes = res["calls"] * 3 * cell_bytes
    log(f"[bench] {ms_step:.1f} ms/step, kernel {1000 * kern_s:.1f} ms, {calls_s:.4e} probe calls/s, "
        f"c3 values {res['calls'] + res['gated']:.4e}, solutions {sorted(got)}")
    ceiling = None
    if not args.no_measure:
        n = 1 << 30
        pms, rms = ctx.measure_gathers(n)
        probe_rate = n / (pms / 1000.0)
        ceiling = {"probe_only_probes_per_s": probe_rate, "raw_random_8B_gathers_per_s": 3 * n / (rms / 1000.0),
                   "kernel_call_rate_frac_of_probe_only": res["calls"] / kern_s / probe_rate}
        log(f"[bench] gather ceiling on this table: {probe_rate:.4e} probes/s; kernel at "
            f"{res['calls'] / kern_s / probe_rate:.1%}")
    strict = None
    if not args.no_strict:
        # the same candidates with the option strict_gate on: same counters but maybe, same solutions
        ctx.set_option("strict_gate", 1)
        try:
            ctx.twin_run(cands)
            skms = []
            t1 = time.perf_counter()
            for _ in range(args.steps):
                sres, shits, ssols = ctx.twin_run(cands)
                skms.append(sres["kernel_ms"])
            swall = time.perf_counter() - t1
        finally:
            ctx.set_option("strict_gate", 0)
        for k in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes"):
            assert sres[k] == res[k], k
        assert {(x["B"], tuple(x["terms"])) for x in ssols} == {(x["B"], tuple(x["terms"])) for x in sols}
        sms = 1000.0 * swall / args.steps
        strict = {"calls_per_s": sres["calls"] / (sms / 1000.0), "ms_per_step": sms, "kernel_ms": float(np.mean(skms)),
                  "speedup_vs_default": ms_step / sms, "queries": sres["calls"] - sres["strict_out"],
                  "strict_out": sres["strict_out"], "maybe": sres["maybe"], "default_maybe": res["maybe"],
                  "solutions_equal": True,
                  "option": "strict_gate=1: the strict residue gate on every key T - c4^6 - c3^6 before its query"}
        log(f"[bench] strict gate: {sms:.1f} ms/step, kernel {strict['kernel_ms']:.1f} ms ({ms_step / sms:.2f}x "
            f"the default); queries {strict['queries']} of {sres['calls']} calls; same solutions")
    plain = None
    if not args.no_plain:
        sub = cands[:: args.twin_plain_every]
        gb, _, _ = ctx.twin_run(sub)
        ctx.set_option("twin_plain", 1)
        t1 = time.perf_counter()
        gp, _, _ = ctx.twin_run(sub)
        tp = time.perf_counter() - t1
        ctx.set_option("twin_plain", 0)
        t1 = time.perf_counter()
        gb2, _, _ = ctx.twin_run(sub)
        tb = time.perf_counter() - t1
        for k in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "maybe"):
            assert gp[k] == gb[k] == gb2[k], k
        plain = {"subset": f"every {args.twin_plain_every}th candidate ({len(sub)} candidates)",
                 "calls_per_s": gp["calls"] / tp, "batched_calls_per_s_same_subset": gb2["calls"] / tb,
                 "speedup_batched_vs_plain": (gb2["calls"] / tb) / (gp["calls"] / tp), "counters_identical": True}
        log(f"[bench] plain form {gp['calls'] / tp:.4e} calls/s on the subset; batched {gb2['calls'] / tb:.4e}; "
            f"batched/plain = {plain['speedup_batched_vs_plain']:.1f}x")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        h, packed = ctx.table_download()
        t = O.Table.from_bytes({f: getattr(h, f) for f, _ in fc5.StoreHeader._fields_}, packed)
        sub = cands[:: args.twin_cpu_every]
        threads = min(16, os.cpu_count() or 1)
        oc, ohits = {}, []
        t1 = t_log = time.perf_counter()
        for k0 in range(0, len(sub), 32):  # in chunks, so that a long sample reports progress
            cc, hh, _ = O.twin_run(sub[k0:k0 + 32].view(O.TWIN_CAND_DTYPE), t, threads=threads)
            oc = {k: oc.get(k, 0) + v for k, v in cc.items()}
            hh = hh.copy()
            hh["cand"] += k0
            ohits.append(hh)
            if time.perf_counter() - t_log > 20:
                t_log = time.perf_counter()
                log(f"[bench]   cpu oracle: {min(k0 + 32, len(sub))}/{len(sub)} candidates, {t_log - t1:.0f}s")
        dt = time.perf_counter() - t1
        gres, ghits, _ = ctx.twin_run(sub)
        for k in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "maybe"):
            assert gres[k] == oc[k], k
        ohits = np.sort(np.concatenate(ohits), order=["cand", "d", "c4", "c3"])
        ghits = np.sort(ghits, order=["cand", "d", "c4", "c3"])
        assert len(ghits) == len(ohits) and all(np.array_equal(ghits[f], ohits[f])
                                                for f in ("cand", "d", "c4", "c3", "fp")), "maybe-hit lists differ"
        what = "all candidates" if args.twin_cpu_every == 1 else f"every {args.twin_cpu_every}-th candidate"
        cpu = {"value": oc["calls"] / dt, "unit": "probe calls/s", "cores": threads, "kind": "port", "host": host_cpu(),
               "sample": f"{what} ({len(sub)} candidates, {oc['calls']} calls, {dt:.1f} s), same table bytes; "
                         f"counters and the {len(ohits)}-entry maybe-hit list identical to the GPU on that sample"}
        log(f"[bench] cpu oracle {oc['calls'] / dt:.4e} calls/s on {len(sub)} candidates, {threads} threads; "
            f"counters and {len(ohits)} maybe-hits identical to the GPU")
    traffic = twin_traffic(lo, hi)
    line = {
        "metric": "617 cls5 twin probe calls/sec, 1 GPU, bit-exact", "value": calls_s,
        "unit": "probe calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (deterministic candidate enumeration of the B range)",
        "config": {"workload": f"(6,1,7) Branch-A class-5 twin, B in [{lo},{hi}] ({len(Bs)} B, {len(cands)} "
                               f"candidates), xor pair table N={N} r={hdr.r} (device-built)",
                   "B_lo": lo, "B_hi": hi, "N": N, "r": int(hdr.r), "candidates": int(len(cands))},
        "eligible_B_per_hour": len(Bs) * 3600.0 / (ms_step / 1000.0),
        "counters": {k: res[k] for k in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes",
                                         "maybe", "solutions")},
        "kernel_ms": float(np.mean(kms)),
        "roofline": {"bound": "hbm", "achieved": alg_bytes / kern_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg_bytes / kern_s / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_twin_batched",
                     "per_unit": f"{3 * cell_bytes} B per probe call (3 cells x {cell_bytes} B)",
                     "traffic_GBps": traffic / kern_s / 1e9 if traffic else None,
                     "traffic_frac": traffic / kern_s / 1e9 / HBM_PEAK_GBS if traffic else None},
        "gather_ceiling": ceiling,
        "strict_gate": strict,
        "plain_form": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def serve_requests(n, seed):
    """Serving workload (--config 5): n band requests, widths log-uniform on [500, 20000] integers,
    starts uniform on [1.0M, 3.5M], classes 4 and 5 (seeded, deterministic)."""
    import random
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        w = int(math.exp(rnd.uniform(math.log(500), math.log(20000))))
        lo = rnd.randint(1000000, 3500000)
        out.append((lo, lo + w - 1))
    return out


def bench_serve(args):
    """Resident serving loop: Poisson arrivals of Stage-1 band requests at a
    fixed fraction of the measured capacity, one shared device table."""
    import random
    seed = 0x615 << 8 | 5
    reqs = serve_requests(args.serve_requests, seed)
    t0 = time.time()
    bands = []
    for lo, hi in reqs:
        cls, Bs, us = fc5.emit_units(lo, hi, classes=(4, 5))
        units = fc5.units_array(Bs, us)
        bands.append((lo, hi, cls, units, fc5.units_calls(units[cls == 5])))
    N = max(83809, max(hi for _, hi in reqs) // 42 + 1)  # the table for B up to 3.52M: 4.32e9 cells > 2^32
    log(f"[bench] serving: {len(reqs)} requests, {sum(len(b[3]) for b in bands)} units, "
        f"B in [{min(r[0] for r in reqs)},{max(r[1] for r in reqs)}], units emitted in {time.time() - t0:.1f}s")
    ctx = fc5.Context(0)
    if args.load_policy is not None:
        ctx.set_option("load_policy", args.load_policy)
    t0 = time.time()
    ctx.table_build(N, args.r, on_gpu=True)
    hdr = ctx.table_info()
    log(f"[bench] device table N={N} cells={hdr.m_cells} (> 2^32: {hdr.m_cells > 2 ** 32}) "
        f"packed={fc5_packed_gb(hdr):.2f} GB built in {time.time() - t0:.1f}s")
    # results alone, for the identity check of a sample of bands
    rnd = random.Random(seed)
    check = sorted(rnd.sample(range(len(bands)), min(args.serve_check, len(bands))))
    alone = {i: ctx.run(bands[i][3], fc5.MODE_BATCHED, cls=bands[i][2]) for i in check}
    ctx.set_option("rows_per_item", 16)
    ctx.serve_start(max_bands=16)

    arrivals = {}

    def run(order, times):
        """submit band i at times[i] seconds after start; returns per-band timings"""
        start = time.perf_counter()
        ids = {}
        pending = list(zip(times, order))
        while pending:
            dt = pending[0][0] - (time.perf_counter() - start)
            if dt > 0:
                time.sleep(min(dt, 0.05))
                continue
            _, i = pending.pop(0)
            ids[i] = ctx.serve_submit(bands[i][3], bands[i][2])
            if not pending:  # the device's busy time over the arrival window
                arrivals["span"] = time.perf_counter() - start
                arrivals["busy_ms"] = ctx.serve_stats()["busy_ms"]
            if len(ids) % 25 == 0:
                st = ctx.serve_stats()
                log(f"[bench]   t={time.perf_counter() - start:.1f}s submitted {st['submitted']} done {st['done']} "
                    f"active {st['active']} pending {st['pending']}")
        for i in order:
            if not ctx.serve_wait(ids[i], 900.0):
                raise SystemExit(f"band {i} not finished after 900 s")
        return ids, time.perf_counter() - start

    # closed-loop capacity: a sample of bands submitted at once
    cal = list(range(min(args.serve_calibrate, len(bands))))
    b0 = ctx.serve_stats()["busy_ms"]
    ids, span = run(cal, [0.0] * len(cal))
    cal_busy = (ctx.serve_stats()["busy_ms"] - b0) / 1000.0 / span
    cal_points = sum(ctx.serve_result(ids[i]).result["calls"] for i in cal)
    for i in cal:
        ctx.serve_release(ids[i])
    cap_points = cal_points / span
    mean_points = float(np.mean([b[4] for b in bands]))  # class-5 closed form per request
    cal_c5 = sum(bands[i][4] for i in cal)
    # capacity in requests/s, scaled from the calibration's class-5 share of its points
    rate = args.serve_load * cap_points * (cal_c5 / cal_points) / mean_points
    log(f"[bench] capacity {cap_points:.4e} grid points/s on {len(cal)} bands at once (device busy "
        f"{cal_busy:.1%} of that run); arrival rate {rate:.3f} requests/s (load {args.serve_load})")
    rnd = random.Random(seed + 1)
    times, t = [], 0.0
    for _ in bands:
        t += rnd.expovariate(rate)
        times.append(t)
    st0 = ctx.serve_stats()
    ids, span = run(list(range(len(bands))), times)
    st1 = ctx.serve_stats()
    lat, points, probes, n_hits = [], 0, 0, 0
    for i in range(len(bands)):
        out = ctx.serve_result(ids[i])
        tm = ctx.serve_timing(ids[i])
        lat.append(tm["t_done"] - tm["t_submit"])
        points += out.result["calls"]
        probes += out.result["probes"]
        n_hits += out.result["n_hits"]
        assert out.by_class[5]["calls"] == bands[i][4], "closed-form class-5 calls"
        r = out.result
        assert r["calls"] == r["neg"] + r["zero"] + r["gated"] + r["probes"] + \
            sum(out.by_class[k]["calls"] - out.by_class[k]["gated"] - out.by_class[k]["probes"] for k in (2, 3, 4))
        if i in alone:
            a = alone[i]
            for k in (2, 3, 4, 5):
                for f in ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe", "jobs", "fail_t"):
                    assert out.by_class[k][f] == a.by_class[k][f], (i, k, f)
            assert np.array_equal(out.hits, a.hits) and out.solutions == a.solutions
    ctx.serve_stop()
    lat = np.array(lat)
    busy_ms = st1["busy_ms"] - st0["busy_ms"]
    n_B = sum(eligible_count(lo, hi) for lo, hi in reqs)
    busy_arr = (arrivals["busy_ms"] - st0["busy_ms"]) / 1000.0 / arrivals["span"]
    log(f"[bench] served {len(bands)} bands in {span:.1f}s: p50 {np.percentile(lat, 50):.2f}s "
        f"p99 {np.percentile(lat, 99):.2f}s, {points / span:.4e} points/s, device busy {busy_ms / 1000 / span:.1%} "
        f"of the whole run, {busy_arr:.1%} over the {arrivals['span']:.1f}s arrival window")
    # roofline of the serving rounds: algorithmic cell bytes of every probe over the summed round time
    cell_bytes = math.ceil(args.r / 8)
    alg = probes * 3 * cell_bytes
    roof = {"bound": "hbm", "achieved": alg / (busy_ms / 1000.0) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": alg / (busy_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, "traffic": None, "kernel": "k_cls5_batched",
            "per_unit": f"{3 * cell_bytes} B per filter probe (3 cells x {cell_bytes} B), over the summed "
                        "event time of the serving rounds"}
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle on a unit sample of the bands checked alone, over the same table bytes
        cu = np.concatenate([bands[i][3] for i in check])
        cc = np.concatenate([bands[i][2] for i in check]).astype(np.uint32)
        every = max(1, len(cu) // 600)
        cpu = cpu_baseline(ctx, cu, cc, every, {"plant_mode": None}, None, np.zeros(0, np.uint64))
        cpu["sample"] = f"units of the {len(check)} bands checked alone: " + cpu["sample"]
    line = {
        "metric": "serving: band latency p50/p99 + throughput, Poisson arrivals, 1 GPU, bit-exact",
        "value": points / span, "unit": "(d,e) grid points + c3 values / s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": 1000.0 * span, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded request stream; deterministic Stage-1 enumeration per band)",
        "config": {"workload": f"serving: {len(bands)} band requests (widths log-uniform [500,20000], starts "
                               f"uniform [1.0M,3.5M], classes 4+5), Poisson arrivals at {args.serve_load} of the "
                               f"measured capacity, one device table N={N} r={hdr.r} ({hdr.m_cells} cells)",
                   "requests": len(bands), "load": args.serve_load, "N": N, "r": int(hdr.r)},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90)),
                      "p99": float(np.percentile(lat, 99)), "max": float(lat.max()), "mean": float(lat.mean())},
        "arrival_rate_per_s": rate, "capacity_points_per_s": cap_points,
        "eligible_B_per_hour": n_B * 3600.0 / span, "filter_probes_per_s": probes / span,
        "device_busy_frac": busy_ms / 1000.0 / span,
        "device_busy_frac_arrival_window": busy_arr, "arrival_window_s": arrivals["span"],
        "device_busy_frac_closed_loop": cal_busy,
        # the offered work as realized by the seeded arrivals, and the busy fraction it implies
        # at the closed-loop rate of busy device time (points per busy second)
        "offered_load_realized": points / (cap_points * arrivals["span"]),
        "device_busy_frac_predicted_from_work": points / (cap_points / cal_busy) / span,
        "rounds": st1["rounds"] - st0["rounds"],
        "bands_checked_alone": len(alone), "n_hits": n_hits,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def table_via_file(ctx, hdr, where):
    """The --load-table path (fourcore_cls5_gpu_v4.cu:568-594, fourcore_xor_store.hpp:147-181):
    write the resident table as a StoreHeader-v2 .bin file, then load it back with
    the exact-N check and the FPR-smoke refusal, which re-uploads and widens the
    cells.  The file is removed afterwards.  Skipped (said loudly, and the
    workload says so) when the directory lacks room for it."""
    import shutil
    import tempfile
    d = where or tempfile.gettempdir()
    path = os.path.join(d, f"fc5_xor_N{hdr.N}_r{hdr.r}_{os.getpid()}.bin")
    need = (hdr.m_cells * hdr.r + 7) // 8 + 8 + 64
    free = shutil.disk_usage(d).free
    if free < need + (2 << 30):
        log(f"[bench] NOT loading the table from a file: {d} has {free / 1e9:.1f} GB free, the file needs "
            f"{need / 1e9:.1f} GB; the device-built table is searched")
        return {"loaded": False, "reason": f"{d}: {free / 1e9:.1f} GB free < {need / 1e9:.1f} GB"}
    before = {f: getattr(hdr, f) for f, _ in fc5.StoreHeader._fields_}
    try:
        t0 = time.time()
        ctx.table_save(path)
        t_save = time.time() - t0
        t0 = time.time()
        ctx.table_load(path, hdr.N)
        t_load = time.time() - t0
    finally:
        if os.path.exists(path):
            os.remove(path)
    after = ctx.table_info()
    assert {f: getattr(after, f) for f, _ in fc5.StoreHeader._fields_} == before, "loaded header differs"
    log(f"[bench] table written to {path} ({need / 1e9:.2f} GB) in {t_save:.1f}s and loaded back in "
        f"{t_load:.1f}s (exact-N check N={hdr.N}, streamed to the device in pinned chunks, widen, FPR smoke)")
    return {"loaded": True, "bytes": need, "save_s": t_save, "load_s": t_load, "dir": d,
            "load": "fc5_table_load: exact-N check, file streamed to the device in 64 MiB pinned chunks, "
                    "cell widening, FPR smoke on the device"}


def hit_T(cls, B, u, d, e):
    """the exact T of a hit: class 5 T(e, d); classes 2-4 T(d) - c3^6 with c3 in e"""
    if cls == 5:
        return (B ** 6 - u ** 6 - (14 * e) ** 6 - (21 * d) ** 6) // 42 ** 6
    f = {2: 14, 3: 21, 4: 7}[cls]
    return (B ** 6 - u ** 6 - (f * d) ** 6) // 42 ** 6 - e ** 6


def strict_leg(ctx, band, base, Bs, us, ucls, args, check, base_value, saved, n_B):
    """The same band with the option strict_gate on (tuned and timed like the
    default leg).  Checks: the gate counters equal the default run's, every
    plant is found, and the hit list is exactly the default hit list restricted
    to hits whose exact T passes the strict gate.  saved: the default leg's
    (rows_per_item, blocks_per_cu), restored afterwards."""
    ctx.set_option("strict_gate", 1)
    ctx.set_option("rows_per_item", 0)  # the variant's own untuned shape
    ctx.set_option("blocks_per_cu", args.blocks_per_cu)
    tune = None
    try:
        if not args.no_autotune:
            tune = band.autotune(sample_every=16)
        check(band.run(fc5.MODE_BATCHED))
        kms, outs = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            outs.append(band.run(fc5.MODE_BATCHED))
            kms.append(outs[-1].result["kernel_ms"])
        wall = time.perf_counter() - t0
    finally:
        ctx.set_option("strict_gate", 0)
        ctx.set_option("rows_per_item", saved[0])
        ctx.set_option("blocks_per_cu", saved[1])
    for o in outs:
        check(o)
    r, r0 = outs[-1].result, base.result
    for k in ("calls", "neg", "zero", "gated", "probes", "jobs", "fail_t"):
        assert r[k] == r0[k], k
    keep = set()
    for h in base.hits:
        i = int(h["unit"])
        cls = int(ucls[i]) if ucls is not None else 5
        if fc5.strict_gate_ok(hit_T(cls, int(Bs[i]), int(us[i]), int(h["d"]), int(h["e"]))):
            keep.add((i, int(h["e"]), int(h["d"]), int(h["flags"]) & 3))
    got = {(int(h["unit"]), int(h["e"]), int(h["d"]), int(h["flags"]) & 3) for h in outs[-1].hits}
    assert got == keep, "strict hits differ from the strict-passing default hits"
    ms = 1000.0 * wall / args.steps
    value = r["calls"] / (ms / 1000.0)
    queries = r["probes"] - r["strict_out"]
    log(f"[bench] strict gate: {ms:.1f} ms/step, kernel {float(np.mean(kms)):.1f} ms, {value:.4e} points/s "
        f"({value / base_value:.2f}x the default leg); queries {queries} of {r['probes']} probes; "
        f"maybe={r['maybe']} side_maybe={r['side_maybe']} (default {r0['maybe']}, {r0['side_maybe']})")
    return {"value": value, "unit": "(d,e) grid points/s", "ms_per_step": ms, "kernel_ms": float(np.mean(kms)),
            "eligible_B_per_hour": n_B * 3600.0 / (ms / 1000.0),
            "speedup_vs_default": value / base_value, "probes": r["probes"], "queries": queries,
            "strict_out": r["strict_out"], "maybe": r["maybe"], "side_maybe": r["side_maybe"],
            "default_maybe": r0["maybe"], "default_side_maybe": r0["side_maybe"],
            "hits_equal_filtered_default": True, "autotune": tune,
            "option": "strict_gate=1: before each xor query T mod 64, 81, 31, 37, 43, 49, 61, 67, 73, 79 must also be sums "
                      "of two sixth powers (a necessary condition for a pair sum that the reference does "
                      "not test); every counter but maybe/side_maybe is the reference's, the hit list is "
                      "the default one without the filter's false positives the strict gate rejects"}


# The reference's published class-5 band times with fourcore_cls5_gpu_v4 on an RTX PRO
# 6000 Blackwell (README.md:146): (B_lo, B_hi, minutes) for 3.5M-4.0M, 4.0M-4.5M,
# 4.5M-5.0M.  Work per B grows as B^3 (units as B, the (d,e) grid as B^2), so each band
# projects to eligible B/hour at another B by (B_mid / B)^3.
REFERENCE_BANDS = ((3500000, 4000000, 49.0), (4000000, 4500000, 71.0), (4500000, 5000000, 98.0))


def reference_projection(cfg, b_per_hour):
    """the reference GPU's eligible class-5 B per hour projected to this band's middle
    from its published band times, and this run's ratio to it (class-5 bands only)"""
    if cfg["classes"] != (5,):
        return None
    mid = (cfg["lo"] + cfg["hi"]) / 2.0
    proj = []
    for lo, hi, minutes in REFERENCE_BANDS:
        per_hour = eligible_count(lo, hi) * 60.0 / minutes
        proj.append(per_hour * (((lo + hi) / 2.0) / mid) ** 3)
    p = sum(proj) / len(proj)
    return {"eligible_B_per_hour": p, "ratio": b_per_hour / p, "derived": True,
            "how": "published band times of fourcore_cls5_gpu_v4 on an RTX PRO 6000 (README.md:146), "
                   "scaled to this band by B^3; a derived figure, not a published one"}


PROBE_MODES = {"auto": -1, "gathers": 0, "partitioned": 1}

# algorithmic bytes per probe of each pipeline stage (records: 8-byte hash,
# then hash + 8-bit partial xor): read + written
PART_STAGES = (("first", "k_part_first", 8 + 8), ("fine0", "k_part_fine", 8 + 8),
               ("lookup0", "k_part_lookup", 8 + 9), ("fine1", "k_part_fine", 9 + 9),
               ("lookup1", "k_part_lookup", 9 + 9), ("fine2", "k_part_fine", 9 + 9),
               ("last", "k_part_last", 9), ("check", "k_part_check", 0))
PART_SLICE_CELLS = 65536  # cells of one third per slice, staged as their low 8 bits


def partition_figures(pst, r, hdr, ms_step):
    """Stage times of the last partitioned run (event-timed, summed over its
    pieces), each stage's algorithmic bytes and rate, and the roofline of the
    dominant kernel: k_part_lookup, two launches per piece."""
    ms, cnt = pst
    pieces = cnt["pieces"]
    probes = r["probes"]
    block = hdr.m_cells // 3
    slice_bytes = -(-block // PART_SLICE_CELLS) * PART_SLICE_CELLS  # one third of the low-bits table
    stages = {}
    for key, kern, per in PART_STAGES:
        b = probes * per + (slice_bytes * pieces if kern in ("k_part_lookup", "k_part_last") else 0)
        stages[key] = {"kernel": kern, "ms": ms[key], "bytes_per_probe": per, "bytes": b,
                       "GBps": b / (ms[key] / 1000.0) / 1e9 if ms[key] > 0 else None}
    lk_ms = (ms["lookup0"] + ms["lookup1"]) / (2 * pieces)
    lk_bytes = (probes * (17 + 18) + 2 * pieces * slice_bytes) / (2 * pieces)
    achieved = lk_bytes / (lk_ms / 1000.0) / 1e9
    pipe_bytes = sum(v["bytes"] for v in stages.values())
    pipe_ms = sum(v["ms"] for v in stages.values())
    log(f"[bench] pipeline stages (ms over {pieces} pieces): "
        + ", ".join(f"{k} {v['ms']:.0f} ({v['GBps'] or 0:.0f} GB/s)" for k, v in stages.items()))
    return {"pieces": pieces, "accepted_hashes": cnt["accepted"], "full_queries": cnt["full_queries"],
            "emit_chunks": cnt["chunks"], "host_ms": {k: ms[k] for k in ("emit", "pipeline", "identify")},
            "stages": stages,
            "pipeline": {"ms": pipe_ms, "bytes": pipe_bytes, "GBps": pipe_bytes / (pipe_ms / 1000.0) / 1e9,
                         "frac": pipe_bytes / (pipe_ms / 1000.0) / 1e9 / HBM_PEAK_GBS},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_part_lookup",
                         "algorithmic_bytes_per_launch": lk_bytes, "avg_launch_ms": lk_ms,
                         "launches_per_step": 2 * pieces,
                         "per_unit": "35 B per probe over a piece's two lookup launches (cell 0: 8 B read, "
                                     "9 B written; cell 1: 9 B read, 9 B written) plus 64 KiB of low cell "
                                     "bits per slice staged in LDS",
                         "traffic_GBps": None, "traffic_frac": None}}


def gathers_leg(ctx, band, base, check, ms_part, cell_bytes, cfg):
    """The same band with three gathers per probe (probe_mode 0, the batched
    kernel alone): its step time, and counters and hit records identical to
    the partitioned run's."""
    ctx.set_option("probe_mode", 0)
    try:
        check(band.run(fc5.MODE_BATCHED))
        t0 = time.perf_counter()
        outs = [band.run(fc5.MODE_BATCHED) for _ in range(2)]
        ms = 1000.0 * (time.perf_counter() - t0) / 2
    finally:
        ctx.set_option("probe_mode", -1)
    for o in outs:
        check(o)
        for k in COUNTER_KEYS:
            assert o.result[k] == base.result[k], k
        ga = np.sort(o.hits, order=["unit", "e", "d", "flags"])
        gb = np.sort(base.hits, order=["unit", "e", "d", "flags"])
        assert np.array_equal(ga, gb), "gathers and partitioned hit lists differ"
    kms = float(np.mean([o.result["kernel_ms"] for o in outs]))
    log(f"[bench] three gathers per probe: {ms:.1f} ms/step (kernel {kms:.1f} ms); partitioned is "
        f"{ms / ms_part:.2f}x faster, counters and hits identical")
    return {"ms_per_step": ms, "kernel_ms": kms, "points_per_s": base.result["calls"] / (ms / 1000.0),
            "speedup_partitioned_vs_gathers": ms / ms_part, "counters_and_hits_identical": True}


def load_profile(name):
    """a committed rocprofv3 counter summary (JSON) under profiles/, or None (said loudly)"""
    p = os.path.join(REPO, "profiles", name)
    if not os.path.exists(p):
        log(f"[bench] no counter summary profiles/{name}: its roofline fields stay null")
        return None
    with open(p) as f:
        return json.load(f)


def measure_ceilings(ctx, hdr, kernel_probe_rate):
    """Roofline microbenchmarks on the resident main table: the probe-only and
    raw random-gather rates, random gathers confined to growing regions of the
    table (L2 / memory-side cache / HBM), a read-only stream over a 24 GiB
    buffer, a copy (calibrates the profiler's byte counters) and the integer
    VALU issue rates."""
    n = 1 << 31
    pms, rms = ctx.measure_gathers(n)
    cms = ctx.measure_copy(4 << 30)
    read_bytes = 24 << 30
    sms = ctx.measure_read(read_bytes)
    probe_rate = n / (pms / 1000.0)
    regions = []
    for mib in (4, 16, 64, 128, 192, 256, 384, 512, 1024, 4096):
        cells = (mib << 20) // 8
        if cells > hdr.m_cells:
            break
        g = 1 << 29
        regions.append({"MiB": mib, "gathers_per_s": g / (ctx.measure_gathers_region(cells, g) / 1000.0)})
    regions.append({"MiB": hdr.m_cells * 8 / 2 ** 20, "gathers_per_s": 3 * n / (rms / 1000.0)})
    valu = {name: ctx.measure_valu(k) for k, name in enumerate(("int32_xad", "mad_u64_u32", "add_u128_carry",
                                                                  "mix_7pct_int64"))}
    ceiling = {"probe_only_probes_per_s": probe_rate, "raw_random_8B_gathers_per_s": 3 * n / (rms / 1000.0),
               "probe_only_line_fill_GBps_at_128B": probe_rate * 3 * 128 / 1e9,
               "copy_GBps": 2 * (4 << 30) / (cms / 1000.0) / 1e9,
               "read_stream_GBps": read_bytes / (sms / 1000.0) / 1e9,
               "random_gathers_by_region": regions,
               "valu_wave_instr_per_s": valu,
               "kernel_probe_rate_frac_of_probe_only": kernel_probe_rate / probe_rate}
    log(f"[bench] gather ceiling: {probe_rate:.4e} probes/s probe-only, "
        f"{3 * n / (rms / 1000.0):.4e} raw gathers/s; kernel at {kernel_probe_rate / probe_rate:.1%}; "
        f"copy {ceiling['copy_GBps']:.0f} GB/s, read stream {ceiling['read_stream_GBps']:.0f} GB/s")
    log("[bench] random gathers by region: " + ", ".join(f"{x['MiB']:.0f} MiB {x['gathers_per_s']:.3e}/s"
                                                        for x in regions))
    log("[bench] VALU wave-instructions/s: " + ", ".join(f"{k} {v:.3e}" for k, v in valu.items()))
    return ceiling


def valu_roofline(prof, ceiling):
    """VALU side of the roofline: the kernel's VALU instruction rate (rocprofv3
    SQ_INSTS_VALU over the launch time of the same run) against the issue rate of
    the microbenchmark with the kernel's INT64 share, and VALUBusy"""
    if not prof:
        return None
    rate = prof["SQ_INSTS_VALU"] / (prof["kernel_ms"] / 1000.0)
    ceil = ceiling["valu_wave_instr_per_s"]["mix_7pct_int64"] if ceiling else None
    return {"busy": prof["VALUBusy"] / 100.0, "achieved": rate, "ceiling": ceil, "unit": "wave instructions/s",
            "frac": rate / ceil if ceil else None, "int64_share": prof["SQ_INSTS_VALU_INT64"] / prof["SQ_INSTS_VALU"],
            "source": prof.get("commands")}


def twin_traffic(lo, hi):
    """HBM bytes per launch from the committed rocprofv3 counter summary of the
    same workload, or None"""
    p = os.path.join(REPO, "profiles", f"traffic_twin617_B{lo}.json")
    return json.load(open(p)).get("traffic_bytes_per_launch") if lo == hi and os.path.exists(p) else None


def fc5_packed_gb(h):
    return (h.m_cells * h.r + 7) // 8 / 1e9


def cpu_baseline(ctx, units, ucls, every, cfg, r, pkeys, offset=0):
    """The oracle (CPU restatement of the reference kernels) timed on every
    k-th unit of the same band, over the same table bytes, on this host; its
    counters and its maybe-hit list are checked equal to the GPU's on that sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    hdr, packed = ctx.table_download()
    t = O.Table.from_bytes({f: getattr(hdr, f) for f, _ in fc5.StoreHeader._fields_}, packed)
    side = O.Table.build_keys(pkeys, hdr.r) if cfg["plant_mode"] == "side" and len(pkeys) else None
    sub, scls = units[offset::every], ucls[offset::every]
    threads = min(16, os.cpu_count() or 1)
    i5 = np.flatnonzero(scls == 5)
    arr, ok = O.pack_units(sub["B"][i5], sub["u"][i5])
    assert ok.all()
    c, hits = {}, []
    t0 = t_log = time.perf_counter()
    chunk = 2048
    for lo in range(0, len(i5), chunk):  # in chunks, so that a long sample reports progress
        n = min(chunk, len(i5) - lo)
        cc, hh, _ = O.peel((O.Unit5 * n)(*arr[lo:lo + n]), n, t, side=side, threads=threads)
        for k, v in cc.items():
            c[k] = c.get(k, 0) + v
        hh = hh.copy()
        hh["unit"] = i5[lo + hh["unit"].astype(np.int64)]
        hits.append(hh)
        if time.perf_counter() - t_log > 20:
            t_log = time.perf_counter()
            log(f"[bench]   cpu oracle: {lo + n}/{len(i5)} class-5 units, {t_log - t0:.0f}s")
    c3 = None
    m3 = np.flatnonzero(scls != 5)
    for lo in range(0, len(m3), chunk):
        m = m3[lo:lo + chunk]
        cc, h3, _ = O.find3(scls[m], sub["B"][m], sub["u"][m], t, threads=threads)
        c3 = cc if c3 is None else {k: c3[k] + v for k, v in cc.items()}
        h3 = h3.copy()
        h3["unit"] = m[h3["unit"].astype(np.int64)]
        hits.append(h3)
        if time.perf_counter() - t_log > 20:
            t_log = time.perf_counter()
            log(f"[bench]   cpu oracle: {lo + len(m)}/{len(m3)} class 2-4 units, {t_log - t0:.0f}s")
    dt = time.perf_counter() - t0
    # the same subset on the GPU, for a like-for-like check
    out = ctx.run(sub, cls=scls)
    for k in ("calls", "neg", "zero", "gated", "probes", "maybe"):
        assert out.by_class[5][k] == c[k], (k, out.by_class[5][k], c[k])
    if c3 is not None:
        for k, ok_ in (("calls", "calls"), ("gated", "gated"), ("probes", "probes"), ("jobs", "jobs"),
                       ("fail_t", "failT")):
            assert sum(out.by_class[cl][k] for cl in (2, 3, 4)) == c3[ok_], k
        c["calls"] += c3["calls"]
    oh = np.sort(np.concatenate(hits) if hits else np.zeros(0, O.HIT_DTYPE), order=["unit", "e", "d", "flags"])
    gh = np.sort(out.hits, order=["unit", "e", "d", "flags"])
    assert len(gh) == len(oh) and all(np.array_equal(gh[f], oh[f]) for f in ("unit", "d", "e", "flags", "fp")), \
        "maybe-hit lists differ"
    log(f"[bench] cpu oracle {c['calls'] / dt:.4e} points/s on {len(sub)} units, {threads} threads, {dt:.1f}s; "
        f"counters and {len(oh)} maybe-hits identical to the GPU")
    return {"value": c["calls"] / dt, "unit": "(d,e) grid points/s", "cores": threads, "kind": "port", "host": host_cpu(),
            "sample": ("all units of the band" if every == 1 else f"every {every}-th unit of the band from unit {offset}")
                      + f" ({len(sub)} units, {c['calls']} grid points, "
                      f"{dt:.1f} s), same table bytes; counters and the {len(oh)}-entry maybe-hit list "
                      f"identical to the GPU on that sample"}


def cpu_reference(config):
    """The reference's own CPU class-5 search (solve_516_v2, every (e, d) with an
    exact pair-index lookup, solve_516_v2.cpp:400-416) as timed by
    oracle/time_ref_cpu.py where the reference sources are present -- this script
    never runs it.  Its figures for the bands of --config 1 and 2 (the solver
    stops at B <= 2.2M), or None."""
    p = os.path.join(REPO, "oracle", "ref_cpu_timing.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    c = t["configs"].get(str(config))
    if c is None:
        return None
    return {"value": c["points_per_s"], "unit": "(d,e) grid points/s", "cores": t["threads"], "kind": "reference",
            "host": t["host"], "where": "timed by oracle/time_ref_cpu.py where the reference sources are, not on this host",
            "sample": f"B in [{c['B_lo']},{c['B_hi']}] ({c['units']} class-5 units, {c['grid_points']} grid points): "
                      f"the reference's solve_516_v2 class-5 search built from its sources, {c['search_s']:.1f} s "
                      f"after its {c['index_build_s']:.1f} s pair-index build; {c['solutions']} solutions"}


if __name__ == "__main__":
    main()
