#!/usr/bin/env python3
"""Time the reference's own CPU class-5 search where its sources are present.

Test infrastructure, run by hand where the reference sources are present and
`make -C oracle ref` has built oracle/_ref/solve_516_v2 from them: the
reference's exact CPU engine (solve_516_v2.cpp:400-416: every (e, d) of a
class-5 unit, an exact lookup of the target in its sorted pair index).
bench.py never runs it: it reports the figures this script writes to
oracle/ref_cpu_timing.json beside its own CPU baseline (the oracle, timed where
the bench runs).

Workloads (the solver's i128 limit is B <= 2.2M), keyed by bench.py's --config:
  1  B in [100001, 102000], every class-5 unit
  2  B in [1000001, 1000100], the first 100 integers of [1000001, 1020000]
The rate is the closed-form grid-point count of the same Stage-1 units over the
search time (the solver's pair-index build, which it reports, is excluded).
"""
import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "six-one-five_amd", "python"))

import fc5  # noqa: E402

WORK = {1: (100001, 102000), 2: (1000001, 1000100)}


def cpu_model():
    with open("/proc/cpuinfo") as f:
        for ln in f:
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    return "unknown"


def main():
    exe = os.path.join(HERE, "_ref", "solve_516_v2")
    if not os.path.exists(exe):
        raise SystemExit("oracle/_ref/solve_516_v2 is missing: run `make -C oracle ref` where the reference is")
    threads = os.cpu_count() or 1
    out = {"host": {"model": cpu_model(), "nproc": threads}, "threads": threads, "configs": {}}
    for cfg, (lo, hi) in WORK.items():
        _, Bs, us = fc5.emit_units(lo, hi)
        points = fc5.units_calls(fc5.units_array(Bs, us))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        r = subprocess.run([exe, str(lo), str(hi), "5"], capture_output=True, text=True, env=env, timeout=1800)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(f"reference solver failed: {r.stderr[-400:]}")
        build_ms = int(re.search(r"index built in (\d+) ms", r.stderr).group(1))
        units = int(re.search(r"unit candidates tested: (\d+)", r.stderr).group(1))
        sols = int(re.search(r"total solutions: (\d+)", r.stderr).group(1))
        assert units == len(Bs) and sols == 0, (units, len(Bs), sols)
        dt = wall - build_ms / 1000.0
        out["configs"][str(cfg)] = {
            "B_lo": lo, "B_hi": hi, "units": units, "grid_points": points, "search_s": dt,
            "index_build_s": build_ms / 1000.0, "points_per_s": points / dt, "solutions": sols}
        print(f"config {cfg}: B in [{lo},{hi}] {units} units {points} points in {dt:.1f} s "
              f"({points / dt:.4e} points/s, {threads} threads)", flush=True)
    with open(os.path.join(HERE, "ref_cpu_timing.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
