"""Writes the golden fixtures in tests/golden/ from the reference pieces that
compile from their own sources (oracle/_ref, built by `make -C oracle ref`).

Run only where the reference sources are present (nowhere else has them):
    python tests/golden/make_golden.py
Outputs (data only: inputs and expected outputs):
    xor_tables.json     SHA-256 and header of reference xor_build_save tables
    units_100001_102000.txt      Stage-1 "cls B u" lines, classes 2-5, B in [100001,102000]
    residues.json       cls5 / cls2-4 free-term residue specs at sampled units
    seeds.json          seed counts per class at sampled B
    twin_cands.json     (6,1,7) class-5 candidate lists (count, SHA-256 of the
                        lines, small lists in full) and the triple-sum sets
    known_617_cls5.json the class-5 rows of the reference's list of known
                        (6,1,7) solutions (617-solutions-clean.txt), with the
                        role decomposition of each
    exact_table.json    the reference's sorted pair index (QuadSumIndex) for
                        small N: entry count, SHA-256 of its "sum i j" lines,
                        the lines in full for N = 40, and equal_sum_range of
                        sampled targets
"""
import hashlib
import json
import os
import random
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
REF_SRC = "/root/reference"


def run(*args):
    return subprocess.run([os.path.join(REF, args[0]), *map(str, args[1:])], check=True,
                          capture_output=True, text=True).stdout


def main():
    if not os.path.exists(os.path.join(REF, "xor_build_save")):
        sys.exit("oracle/_ref not built: make -C oracle ref")
    tables = []
    with tempfile.TemporaryDirectory() as td:
        for N, r in [(40, 48), (300, 48), (2000, 16), (2428, 48), (2428, 16), (2428, 8)]:
            p = os.path.join(td, f"t_{N}_{r}.bin")
            run("xor_build_save", "--N", N, "--r", r, "--out", p)
            b = open(p, "rb").read()
            hdr = b[:64]
            tables.append({"N": N, "r": r, "bytes": len(b), "sha256": hashlib.sha256(b).hexdigest(),
                           "header_hex": hdr.hex()})
    json.dump(tables, open(os.path.join(HERE, "xor_tables.json"), "w"), indent=1)

    open(os.path.join(HERE, "units_100001_102000.txt"), "w").write(run("ref_classes", "units", 100001, 102000, "2345"))

    rnd = random.Random(615)
    res = []
    # sampled cls5 units across the configurations' B ranges (pre and post 2^128)
    for lo, hi in [(100001, 102000), (1000001, 1020000), (2000001, 2010000), (3200001, 3210000),
                   (3500000, 3520000)]:
        lines = run("ref_classes", "units", lo, lo + 300, "45").split("\n")
        lines = [l for l in lines if l]
        for l in rnd.sample(lines, 8):
            cls, B, u = map(int, l.split())
            if cls == 5:
                out = run("ref_classes", "cls5", B, u).split("\n")
                e = list(map(int, out[0].split()[1:]))
                d = list(map(int, out[1].split()[1:]))
                res.append({"cls": 5, "B": B, "u": u, "e_mod": e[0], "e_fmax": e[1], "e_res": e[2:],
                            "d_mod": d[0], "d_fmax": d[1], "d_res": d[2:]})
            else:
                out = run("ref_classes", "cls234", B, u, 4).split()
                res.append({"cls": 4, "B": B, "u": u, "d_mod": int(out[1]), "d_fmax": int(out[2]),
                            "d_res": list(map(int, out[3:]))})
        for cls in (2, 3):
            lines = [l for l in run("ref_classes", "units", lo, lo + 2000, str(cls)).split("\n") if l]
            for l in rnd.sample(lines, min(4, len(lines))):
                c, B, u = map(int, l.split())
                out = run("ref_classes", "cls234", B, u, c).split()
                res.append({"cls": c, "B": B, "u": u, "d_mod": int(out[1]), "d_fmax": int(out[2]),
                            "d_res": list(map(int, out[3:]))})
    json.dump(res, open(os.path.join(HERE, "residues.json"), "w"), indent=0)

    seeds = []
    for B in [100003, 1000003, 2000003, 3200003, 3519997] + rnd.sample(range(100001, 3520000), 20):
        if B % 2 == 0 or B % 3 == 0 or B % 7 == 0:
            continue
        seeds.append({"B": B, "counts": list(map(int, run("ref_classes", "seeds", B).split()))})
    json.dump(seeds, open(os.path.join(HERE, "seeds.json"), "w"), indent=0)
    known = known_617_cls5(os.path.join(REF_SRC, "617-solutions-clean.txt"))
    json.dump(known, open(os.path.join(HERE, "known_617_cls5.json"), "w"), indent=0)

    twin = {"tri": list(map(int, run("ref_617", "tri").split())), "cands": []}
    Bs = sorted({k["B"] for k in known} | {1009, 2003, 5003, 10007, 100003, 1000003, 2000003, 2353001, 2353003})
    for B, band in [(B, (0.0, 1.0)) for B in Bs] + [(400471, (0.3, 0.7)), (100003, (0.5, 1.0))]:
        text = run("ref_617", "cands", B, band[0], band[1])
        lines = [l for l in text.split("\n") if l]
        rec = {"B": B, "u_band": list(band), "count": len(lines),
               "sha256": hashlib.sha256("\n".join(lines).encode()).hexdigest()}
        if len(lines) <= 40:
            rec["lines"] = lines
        twin["cands"].append(rec)
    json.dump(twin, open(os.path.join(HERE, "twin_cands.json"), "w"), indent=0)

    exact = []
    for N in (40, 300, 1000):
        lines = [l for l in run("ref_quad", "table", N).split("\n") if l]
        rec = {"N": N, "count": len(lines), "sha256": hashlib.sha256("\n".join(lines).encode()).hexdigest()}
        if N == 40:
            rec["lines"] = lines
        if N == 300:
            sums = [int(l.split()[0]) for l in lines]
            targets = [sums[0], sums[-1]] + rnd.sample(sums, 24) + [x + 1 for x in rnd.sample(sums, 8)] + [1, 3]
            rec["ranges"] = [list(map(int, l.split())) for l in run("ref_quad", "range", N, *targets).split("\n") if l]
        exact.append(rec)
    json.dump(exact, open(os.path.join(HERE, "exact_table.json"), "w"), indent=0)
    print("golden fixtures written")


def known_617_cls5(path):
    """Class-5 rows (the odd, not-divisible-by-3 and not-divisible-by-7 terms are
    three different terms) of the known-solution list, 2nd kind (7 does not divide B)."""
    out = []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        m = re.match(r"\s*(\d+)\s*=\s*([\d\s+]+)\|", line)
        B = int(m.group(1))
        t = sorted(int(x) for x in m.group(2).split("+"))
        assert sum(x ** 6 for x in t) == B ** 6, line
        if B % 7 == 0:
            continue
        odd = [x for x in t if x % 2]
        n3 = [x for x in t if x % 3]
        n7 = [x for x in t if x % 7]
        if len(odd) != 1 or len(n3) != 1 or len(n7) != 1 or len({odd[0], n3[0], n7[0]}) != 3:
            continue
        cores = sorted(x // 42 for x in t if x not in (odd[0], n3[0], n7[0]))
        out.append({"B": B, "terms": t, "d": odd[0] // 21, "u": n3[0], "w": n7[0], "cores": cores})
    return out


if __name__ == "__main__":
    main()
