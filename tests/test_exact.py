"""The exact sorted pair table on the CPU: the oracle's restatement of the
reference's sorted pair index (QuadSumIndex, quad_sum.hpp:8-45) against the
golden fixture written from the reference's own index (oracle/_ref/ref_quad),
and against the reference binary itself where it is built."""
import hashlib
import json
import os
import random
import subprocess

import pytest

import oracle as O
from conftest import GOLDEN, REF_DIR

GOLD = json.load(open(os.path.join(GOLDEN, "exact_table.json")))


def lines_of(N, extra=()):
    ent, s = O.exact_table(N, extra)
    return [f"{v} {a} {b}" for v, a, b in (O.exact_decode(e, s) for e in ent)], ent, s


@pytest.mark.parametrize("rec", GOLD, ids=lambda r: f"N{r['N']}")
def test_oracle_table_matches_reference_index(rec):
    lines, ent, s = lines_of(rec["N"])
    assert len(lines) == rec["count"] == rec["N"] * (rec["N"] + 1) // 2
    if "lines" in rec:
        assert lines == rec["lines"]
    assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == rec["sha256"]
    assert s == rec["N"].bit_length()
    for lo, hi in ((0, 1), (len(ent) - 1, len(ent))):
        v, a, b = O.exact_decode(ent[lo], s)
        assert 1 <= a <= b <= rec["N"] and a ** 6 + b ** 6 == v


def test_oracle_equal_ranges_match_reference():
    rec = next(r for r in GOLD if "ranges" in r)
    ent, s = O.exact_table(rec["N"])
    for T, lo, hi in rec["ranges"]:
        assert O.exact_range(ent, s, T) == (lo, hi), T


def test_extra_values_sit_among_the_sums():
    rnd = random.Random(7)
    N = 60
    sums = sorted(a ** 6 + b ** 6 for a in range(1, N + 1) for b in range(a, N + 1))
    extra = [sums[5], sums[100] + 1, 1, 2 * N ** 6 + 17] + [rnd.randrange(1, 2 * N ** 6) for _ in range(20)]
    ent, s = O.exact_table(N, extra)
    assert len(ent) == len(sums) + len(extra)
    assert [e >> s for e in ent] == sorted(sums + extra)
    # an extra equal to a pair sum sits beside it, first (a = 0)
    lo, hi = O.exact_range(ent, s, sums[5])
    assert hi - lo == 2 and O.exact_decode(ent[lo], s)[1] == 0


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "ref_quad")), reason="oracle/_ref not built")
def test_oracle_matches_reference_binary_n500():
    out = subprocess.run([os.path.join(REF_DIR, "ref_quad"), "table", "500"], capture_output=True, text=True,
                         check=True).stdout.split("\n")[:-1]
    assert lines_of(500)[0] == out


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "solve_516_v2")), reason="oracle/_ref not built")
def test_reference_cpu_solver_agrees_on_units():
    """the reference's own CPU solver (the CPU baseline below 2.2M) passes its
    self-test and searches the same class-5 units as Stage 1 here, finding no
    solution, on B in [100001, 102000]"""
    import fc5

    exe = os.path.join(REF_DIR, "solve_516_v2")
    assert subprocess.run([exe, "--selftest"], capture_output=True, text=True).returncode == 0
    r = subprocess.run([exe, "100001", "102000", "5"], capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="4"), check=True)
    _, Bs, _ = fc5.emit_units(100001, 102000)
    assert f"unit candidates tested: {len(Bs)}" in r.stderr and "total solutions: 0" in r.stderr
