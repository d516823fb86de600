"""Writes the golden fixtures in tests/golden/ from the reference pieces that
compile from their own sources (oracle/_ref, built by `make -C oracle ref`).

Run in the build container only (the reference is not present elsewhere):
    python tests/golden/make_golden.py
Outputs (data only: inputs and expected outputs):
    xor_tables.json     SHA-256 and header of reference xor_build_save tables
    units_cfg1.txt      Stage-1 "cls B u" lines, classes 2-5, B in [100001,102000]
    residues.json       cls5 / cls2-4 free-term residue specs at sampled units
    seeds.json          seed counts per class at sampled B
"""
import hashlib
import json
import os
import random
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")


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

    open(os.path.join(HERE, "units_cfg1.txt"), "w").write(run("ref_classes", "units", 100001, 102000, "2345"))

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
    print("golden fixtures written")


if __name__ == "__main__":
    main()
