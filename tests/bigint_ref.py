"""Independent Python big-integer statements used to pin the oracle.

These are written from the mathematics of the class-5 peel, not from the
oracle's code: exact residuals with Python integers, the residue classes as
brute-force sixth-root searches, the gate sets as brute-force pair sums.
"""
M42 = 42 ** 6


def pair_sum_set(m):
    p = [pow(a, 6, m) for a in range(m)]
    return {(x + y) % m for x in p for y in p}


S504 = pair_sum_set(504)
S247 = pair_sum_set(247)
# the strict gate (an option of this build): the same condition modulo 64, 81, 31, 37, 43,
# 49, 61, 67, 73, 79
STRICT = {m: pair_sum_set(m) for m in (64, 81, 31, 37, 43, 49, 61, 67, 73, 79)}


def strict_ok(T):
    return all(T % m in s for m, s in STRICT.items())


def residues(t, m):
    """units a in [1,m) with a^6 = t (mod m)"""
    from math import gcd
    return [a for a in range(1, m) if gcd(a, m) == 1 and pow(a, 6, m) == t % m]


def cls5_grid(B, u, fmax1=None, fmax2=None):
    te = pow(B, 6, 729) * pow(pow(14, 6, 729), -1, 729) % 729
    td = (pow(B, 6, 64) - pow(u, 6, 64)) * pow(pow(21, 6, 64), -1, 64) % 64
    f1 = (B - 1) // 14 if fmax1 is None else fmax1
    f2 = (B - 1) // 21 if fmax2 is None else fmax2
    es = sorted(e for r in residues(te, 729) for e in range(r, f1 + 1, 729) if e)
    ds = sorted(d for r in residues(td, 64) for d in range(r, f2 + 1, 64) if d)
    return es, ds


def peel_counts(B, u, es, ds, contains=None, use_gate=True, strict=False):
    """counters (calls, neg, zero, gated, probes, maybe, queries) and maybe-hit
    list; strict: the strict gate before every query"""
    c = dict(calls=0, neg=0, zero=0, gated=0, probes=0, maybe=0, queries=0)
    hits = []
    base = B ** 6 - u ** 6
    d6 = [(21 * d) ** 6 for d in ds]
    for e in es:
        X = base - (14 * e) ** 6
        for d, D in zip(ds, d6):
            c["calls"] += 1
            R = X - D
            if R < 0:
                c["neg"] += 1
                continue
            if R == 0:
                c["zero"] += 1
                continue
            assert R % M42 == 0
            T = R // M42
            if use_gate and not (T % 504 in S504 and T % 247 in S247):
                c["gated"] += 1
                continue
            c["probes"] += 1
            if strict:
                if not strict_ok(T):
                    continue
                c["queries"] += 1
            fp = T % 2 ** 64
            if contains is not None and contains(fp):
                c["maybe"] += 1
                hits.append((e, d, fp))
    return c, hits


MASK64 = 2 ** 64 - 1


def mix64(x):
    x &= MASK64
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & MASK64
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & MASK64
    x ^= x >> 33
    return x


def xor_query(header, packed, key):
    """membership query over a StoreHeader-v2 packed table, from the format"""
    r, m, seed, n = header["r"], header["m_cells"], header["mix_seed"], header["n_keys"]
    if n == 0 or m < 3:
        return False
    block = m // 3
    mask = (1 << r) - 1
    h = mix64(key + seed)
    cells = [h % block, mix64(h ^ 0x9E3779B97F4A7C15) % block + block,
             mix64(h ^ 0xBF58476D1CE4E5B9) % block + 2 * block]
    got = 0
    for c in cells:
        bit = c * r
        word = int.from_bytes(packed[bit // 8: bit // 8 + 16], "little")
        got ^= (word >> (bit % 8)) & mask
    return got == (h & mask)


F_OF = {2: 14, 3: 21, 4: 7}
MOD_OF = {2: 729, 3: 64, 4: 46656}


def cls234_residues(cls, B, u):
    """free-term residues from the definition: units d mod `mod` with
    (f d)^6 = B^6 - u^6 (mod 2^6 for cls3 and cls4, mod 3^6 for cls2 and cls4)"""
    from math import gcd
    f, m = F_OF[cls], MOD_OF[cls]
    rhs = (pow(B, 6, m) - pow(u, 6, m)) % m
    return [d for d in range(1, m) if gcd(d, m) == 1 and pow(f * d, 6, m) == rhs]


def iroot6(n):
    a = int(round(n ** (1.0 / 6.0))) if n > 0 else 0
    while (a + 1) ** 6 <= n:
        a += 1
    while a > 0 and a ** 6 > n:
        a -= 1
    return a


def find3_counts(cls, B, u, contains=None, use_gate=True, strict=False):
    f, m = F_OF[cls], MOD_OF[cls]
    fmax = (B - 1) // f
    lim = (B - 1) // 42
    c = dict(jobs=0, failT=0, calls=0, gated=0, probes=0, maybe=0, queries=0)
    hits = []
    base = B ** 6 - u ** 6
    for r in cls234_residues(cls, B, u):
        for d in range(r, fmax + 1, m):
            R = base - (f * d) ** 6
            if R <= 0 or R % M42:
                c["failT"] += 1
                continue
            T = R // M42
            c["jobs"] += 1
            hi = min(iroot6(T), lim)
            lo = 1
            while lo <= hi and 3 * lo ** 6 < T:
                lo += 1
            for c3 in range(lo, hi + 1):
                c["calls"] += 1
                V = T - c3 ** 6
                if use_gate and not (V % 504 in S504 and V % 247 in S247):
                    c["gated"] += 1
                    continue
                c["probes"] += 1
                if strict:
                    if not strict_ok(V):
                        continue
                    c["queries"] += 1
                fp = V % 2 ** 64
                if contains is not None and contains(fp):
                    c["maybe"] += 1
                    hits.append((d, c3, fp))
    return c, hits


# ---- (6,1,7) class-5 twin ---------------------------------------------------
def triple_sum_set(m):
    """residues of x^6 + y^6 + z^6 mod m over all x, y, z"""
    p = {pow(a, 6, m) for a in range(m)}
    p2 = {(x + y) % m for x in p for y in p}
    return {(x + y) % m for x in p2 for y in p}


T504 = triple_sum_set(504)
T247 = triple_sum_set(247)


def twin_counts(cand, contains=None, c4_div=3, use_gate=True, use_tri=True, tile=0, zmax=1, strict=False):
    """k_find4_cls5 from its definition: every grid d, every c4 of the window,
    every c3 of the inner window; Python integers throughout.  tile = 0 visits
    each c4 once; tile > 0 runs the reference launch's zmax c4 tiles, tile z
    from lo4 + z * tile to hi4 (no tile end)."""
    base = cand["base"]
    lim, fmax1 = cand["lim"], cand["fmax1"]
    c = dict(jobs=0, c4_total=0, c4_tri_skipped=0, calls=0, gated=0, probes=0, maybe=0, queries=0)
    hits = []
    ds = sorted(d for r in cand["res"] for d in range(r, fmax1 + 1, 64) if d)
    for d in ds:
        R = base - (21 * d) ** 6
        if R <= 0:
            continue
        assert R % M42 == 0
        T = R // M42
        hi4 = min(iroot6(T), lim)
        if hi4 < 1:
            continue
        lo4 = 1
        while c4_div * lo4 ** 6 < T:
            lo4 += 1
        if lo4 > hi4:
            continue
        c["jobs"] += 1
        starts = [lo4 + z * tile for z in range(zmax)] if tile else [lo4]
        for c4 in (c4 for s in starts for c4 in range(s, hi4 + 1)):
            T4 = T - c4 ** 6
            if use_tri:
                c["c4_total"] += 1
                if not (T4 % 504 in T504 and T4 % 247 in T247):
                    c["c4_tri_skipped"] += 1
                    continue
            if T4 <= 0:
                continue
            hi3 = min(iroot6(T4), lim, c4)
            if hi3 < 1:
                continue
            lo3 = 1
            while 3 * lo3 ** 6 < T4:
                lo3 += 1
            for c3 in range(lo3, hi3 + 1):
                V = T4 - c3 ** 6
                if use_gate and not (V % 504 in S504 and V % 247 in S247):
                    c["gated"] += 1
                    continue
                c["calls"] += 1
                c["probes"] += 3
                if strict:
                    if not strict_ok(V):
                        continue
                    c["queries"] += 1
                fp = V % 2 ** 64
                if contains is not None and contains(fp):
                    c["maybe"] += 1
                    hits.append((d, c4, c3, fp))
    return c, hits
