"""ctypes binding of the oracle (oracle/libfc5_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check the product's results.  It restates the
reference algorithm on the CPU (see fc5_oracle.c for the file:line map).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfc5_oracle.so")
_lib = C.CDLL(LIB_PATH)
_P = C.c_void_p


class Header(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("magic", "version", "kind", "r", "w", "N")] + \
               [(n, C.c_uint64) for n in ("n_keys", "m_cells", "mix_seed")] + \
               [(n, C.c_uint32) for n in ("shard_count", "shard_index", "shard_mode", "reserved")]


class Xor(C.Structure):
    _fields_ = [("hdr", Header), ("packed", C.POINTER(C.c_uint8)), ("packed_len", C.c_size_t)]


class Unit5(C.Structure):
    _fields_ = [("B", C.c_uint64), ("u", C.c_uint64), ("base", C.c_uint64 * 3)] + \
               [(n, C.c_uint32) for n in ("lim", "mod1", "mod2", "fmax1", "fmax2", "nres1", "nres2")] + \
               [("res", C.c_uint32 * 24)]


class Counts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe",
                                          "queries")]


class Hit(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("unit", "d", "e", "flags")] + [("fp", C.c_uint64)]


_lib.or_emit_units.restype = C.c_int64
_lib.or_emit_units.argtypes = [C.c_int64, C.c_int64, C.c_uint, C.c_double, C.c_double, _P, _P, _P, C.c_int64]
_lib.or_mod192.restype = C.c_uint64
_lib.or_mod192.argtypes = [_P, C.c_uint64, C.c_uint64]
_lib.or_mix64.restype = C.c_uint64
_lib.or_mix64.argtypes = [C.c_uint64]
_lib.or_inv216.restype = C.c_uint64
_lib.or_build_pairs.argtypes = [C.c_int, C.c_uint32, _P, C.c_size_t, C.POINTER(Xor)]
_lib.or_build_xor.argtypes = [_P, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int, C.POINTER(Xor)]
_lib.or_xor_load.argtypes = [C.c_char_p, C.POINTER(Xor), C.c_int]
_lib.or_xor_save.argtypes = [C.c_char_p, C.POINTER(Xor)]
_lib.or_xor_free.argtypes = [C.POINTER(Xor)]
_lib.or_xor_might_contain.argtypes = [C.POINTER(Xor), C.c_uint64]
_lib.or_fpr_smoke.argtypes = [C.POINTER(Xor), C.c_int]
_lib.or_pack_unit5.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(Unit5)]
_lib.or_cls5_peel.argtypes = [_P, C.c_int, C.POINTER(Xor), C.POINTER(Xor), C.c_int, C.POINTER(Counts), _P,
                              C.c_uint64, C.POINTER(C.c_uint64), C.c_int]
_lib.or_T_peel2.argtypes = [C.c_uint64] * 4 + [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
_lib.or_verify_hit.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                               _P, _P, C.c_int]
_lib.or_seeds_for_B.argtypes = [C.c_int64, C.c_int, _P, C.c_int]
_lib.or_free_de_cls5.argtypes = [C.c_int64, C.c_int64, _P, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), _P, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]
_lib.or_pow6_192.argtypes = [C.c_uint32, _P]
class F3Counts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("jobs", "failT", "calls", "gated", "probes", "maybe", "queries")]


_lib.or_find3.argtypes = [_P, _P, _P, C.c_int, C.POINTER(Xor), C.c_int, C.POINTER(F3Counts), _P, C.c_uint64,
                          C.POINTER(C.c_uint64), C.c_int]
_lib.or_T_peel1.argtypes = [C.c_uint64] * 4 + [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
_lib.or_verify_find3.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                                 _P, _P, C.c_int]
_lib.or_free_d_cls234.argtypes = [C.c_int64, C.c_int64, C.c_int, _P, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]
_lib.or_splitmix64.restype = C.c_uint64
_lib.or_splitmix64.argtypes = [C.POINTER(C.c_uint64)]

HIT_DTYPE = np.dtype([("unit", "<u4"), ("d", "<u4"), ("e", "<u4"), ("flags", "<u4"), ("fp", "<u8")])


def emit_units(lo, hi, classes=(5,), u_band=(0.0, 1.0)):
    mask = 0
    for c in classes:
        mask |= 1 << c
    n = _lib.or_emit_units(lo, hi, mask, u_band[0], u_band[1], None, None, None, 0)
    cls = np.zeros(n, np.int64)
    Bs = np.zeros(n, np.int64)
    us = np.zeros(n, np.int64)
    _lib.or_emit_units(lo, hi, mask, u_band[0], u_band[1], cls.ctypes.data_as(_P), Bs.ctypes.data_as(_P),
                       us.ctypes.data_as(_P), n)
    return cls, Bs, us


def seeds(B, cls):
    out = np.zeros(4096, np.int64)
    n = _lib.or_seeds_for_B(B, cls, out.ctypes.data_as(_P), 4096)
    return out[:n].tolist()


def free_de_cls5(B, u):
    er = np.zeros(16, np.int64)
    dr = np.zeros(16, np.int64)
    ne, nd = C.c_int(), C.c_int()
    em, ef, dm, df = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    _lib.or_free_de_cls5(B, u, er.ctypes.data_as(_P), C.byref(ne), C.byref(em), C.byref(ef),
                         dr.ctypes.data_as(_P), C.byref(nd), C.byref(dm), C.byref(df))
    return (er[:ne.value].tolist(), em.value, ef.value), (dr[:nd.value].tolist(), dm.value, df.value)


def pow6_192(x):
    o = np.zeros(3, np.uint64)
    _lib.or_pow6_192(x, o.ctypes.data_as(_P))
    return int(o[0]) | int(o[1]) << 64 | int(o[2]) << 128


def mod192(v, m, c):
    a = np.array([v & (2**64 - 1), (v >> 64) & (2**64 - 1), v >> 128], np.uint64)
    return _lib.or_mod192(a.ctypes.data_as(_P), m, c)


def mix64(x):
    return _lib.or_mix64(x)


def inv216():
    return _lib.or_inv216()


def splitmix64_stream(seed, n):
    s = C.c_uint64(seed)
    return [_lib.or_splitmix64(C.byref(s)) for _ in range(n)]


def T_peel2(B, u, d, e):
    lo, hi = C.c_uint64(), C.c_uint64()
    if not _lib.or_T_peel2(B, u, d, e, C.byref(lo), C.byref(hi)):
        return None
    return lo.value | hi.value << 64


class Table:
    """An oracle xor table (owned C memory)."""

    def __init__(self):
        self.x = Xor()

    @classmethod
    def build_pairs(cls, N, r=48, extra=None):
        t = cls()
        ex = np.ascontiguousarray(extra if extra is not None else np.zeros(0, np.uint64), dtype=np.uint64)
        rc = _lib.or_build_pairs(N, r, ex.ctypes.data_as(_P) if len(ex) else None, len(ex), C.byref(t.x))
        if rc:
            raise RuntimeError("oracle build stalled")
        return t

    @classmethod
    def build_keys(cls, keys, r=48, base_seed=0x615615615615615, tries=64):
        t = cls()
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        rc = _lib.or_build_xor(k.ctypes.data_as(_P), len(k), r, base_seed, tries, C.byref(t.x))
        if rc:
            raise RuntimeError("oracle build stalled")
        return t

    @classmethod
    def from_bytes(cls, header_fields, packed):
        """wrap a table held in memory (header field dict, uint8 array); the
        array is kept alive by the Table and never freed by the oracle"""
        t = cls()
        for f, _ in Header._fields_:
            setattr(t.x.hdr, f, int(header_fields[f]))
        t._keep = packed
        t.x.packed = packed.ctypes.data_as(C.POINTER(C.c_uint8))
        t.x.packed_len = len(packed)
        t._borrowed = True
        return t

    @classmethod
    def load(cls, path, expected_N=0):
        t = cls()
        rc = _lib.or_xor_load(path.encode(), C.byref(t.x), expected_N)
        if rc:
            raise OSError(rc, f"oracle load {path}")
        return t

    def save(self, path):
        if _lib.or_xor_save(path.encode(), C.byref(self.x)):
            raise OSError(path)

    def header(self):
        h = self.x.hdr
        return {f: getattr(h, f) for f, _ in Header._fields_}

    def header_struct(self):
        return self.x.hdr

    def packed(self):
        return bytes(C.string_at(self.x.packed, self.x.packed_len)) if self.x.packed_len else b""

    def contains(self, key):
        return bool(_lib.or_xor_might_contain(C.byref(self.x), key))

    def fpr_smoke(self, trials=200000):
        return _lib.or_fpr_smoke(C.byref(self.x), trials)

    def __del__(self):
        if getattr(self, "_borrowed", False):
            return
        try:
            _lib.or_xor_free(C.byref(self.x))
        except Exception:
            pass


def pack_units(B, u):
    arr = (Unit5 * len(B))()
    ok = np.zeros(len(B), bool)
    for i, (b, uu) in enumerate(zip(B, u)):
        ok[i] = bool(_lib.or_pack_unit5(int(b), int(uu), C.byref(arr[i])))
    return arr, ok


def peel(units_arr, n, table, side=None, use_gate=True, threads=8, cap=1 << 20, strict=False):
    """k_cls5_192 semantics over packed units: counters dict and sorted hits.
    strict: the strict residue gate before every query (an option of this build)."""
    c = Counts()
    hits = np.zeros(cap, HIT_DTYPE)
    nh = C.c_uint64()
    _lib.or_cls5_peel(C.cast(units_arr, _P), n, C.byref(table.x) if table is not None else None,
                      C.byref(side.x) if side is not None else None, (1 if use_gate else 0) | (2 if strict else 0),
                      C.byref(c),
                      hits.ctypes.data_as(_P), cap, C.byref(nh), threads)
    hits = hits[:min(nh.value, cap)]
    hits = np.sort(hits, order=["unit", "e", "d", "flags"])
    return {f: getattr(c, f) for f, _ in Counts._fields_}, hits, nh.value


def verify_hit(B, u, lim, d, e, fp, N, cap=16):
    terms = np.zeros(cap * 5, np.int64)
    g = np.zeros(cap, np.int64)
    n = _lib.or_verify_hit(B, u, lim, d, e, fp, N, terms.ctypes.data_as(_P), g.ctypes.data_as(_P), cap)
    return [(terms[5 * i:5 * i + 5].tolist(), int(g[i])) for i in range(min(n, cap))]


def find3(cls, B, u, table, use_gate=True, threads=8, cap=1 << 20, strict=False):
    """classes 2-4 find3 semantics: counters dict and sorted hits (e holds c3);
    strict: the strict residue gate before every query"""
    cl = np.ascontiguousarray(cls, dtype=np.int32)
    Bv = np.ascontiguousarray(B, dtype=np.uint64)
    uv = np.ascontiguousarray(u, dtype=np.uint64)
    c = F3Counts()
    hits = np.zeros(cap, HIT_DTYPE)
    nh = C.c_uint64()
    _lib.or_find3(cl.ctypes.data_as(_P), Bv.ctypes.data_as(_P), uv.ctypes.data_as(_P), len(cl),
                  C.byref(table.x) if table is not None else None, (1 if use_gate else 0) | (2 if strict else 0),
                  C.byref(c),
                  hits.ctypes.data_as(_P), cap, C.byref(nh), threads)
    hits = np.sort(hits[:min(nh.value, cap)], order=["unit", "e", "d", "flags"])
    return {f: getattr(c, f) for f, _ in F3Counts._fields_}, hits, nh.value


def T_peel1(B, u, f, d):
    lo, hi = C.c_uint64(), C.c_uint64()
    if not _lib.or_T_peel1(B, u, f, d, C.byref(lo), C.byref(hi)):
        return None
    return lo.value | hi.value << 64


def free_d_cls234(B, u, cls):
    r = np.zeros(64, np.int64)
    n, m, f = C.c_int(), C.c_int64(), C.c_int64()
    _lib.or_free_d_cls234(B, u, cls, r.ctypes.data_as(_P), C.byref(n), C.byref(m), C.byref(f))
    return r[:n.value].tolist(), m.value, f.value


def verify_find3(cls, B, u, d, c, fp, N, cap=16):
    terms = np.zeros(cap * 5, np.int64)
    g = np.zeros(cap, np.int64)
    n = _lib.or_verify_find3(cls, B, u, d, c, fp, N, terms.ctypes.data_as(_P), g.ctypes.data_as(_P), cap)
    return [(terms[5 * i:5 * i + 5].tolist(), int(g[i])) for i in range(min(n, cap))]


# ---- (6,1,7) Branch-A class-5 twin ---------------------------------------
TWIN_CAND_DTYPE = np.dtype([("B", "<u8"), ("u", "<u8"), ("w", "<u8"), ("base_lo", "<u8"), ("base_hi", "<u8"),
                            ("lim", "<u4"), ("fmax1", "<u4"), ("mod1", "<u4"), ("nres1", "<u4"),
                            ("res", "<u4", (8,))])
TWIN_HIT_DTYPE = np.dtype([("cand", "<u4"), ("d", "<u4"), ("c4", "<u4"), ("c3", "<u4"), ("fp", "<u8")])
assert TWIN_CAND_DTYPE.itemsize == 88 and TWIN_HIT_DTYPE.itemsize == 24


class TwinCounts(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes",
                                          "maybe", "queries")]


_lib.or_twin_cands.argtypes = [C.c_int64, C.c_double, C.c_double, _P, C.c_int]
_lib.or_build_tri.argtypes = [_P, _P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
_lib.or_twin_run.argtypes = [_P, C.c_int, C.POINTER(Xor), C.c_int, C.c_int, C.c_int, C.POINTER(TwinCounts), _P,
                             C.c_uint64, C.POINTER(C.c_uint64), C.c_int, C.c_int]
_lib.or_twin_verify.argtypes = [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, _P, _P, C.c_int]


def twin_cands(B, u_band=(0.0, 1.0)):
    """gen_class_cands cls5 of solve_617_v1 for one B (structured array)"""
    n = _lib.or_twin_cands(B, u_band[0], u_band[1], None, 0)
    out = np.zeros(max(n, 1), TWIN_CAND_DTYPE)
    _lib.or_twin_cands(B, u_band[0], u_band[1], out.ctypes.data_as(_P), n)
    return out[:n]


def twin_cand(B, u, w, base, lim, fmax1, res, mod1=64):
    """one hand-made candidate (plants and fixtures)"""
    c = np.zeros(1, TWIN_CAND_DTYPE)
    c["B"], c["u"], c["w"] = B, u, w
    c["base_lo"], c["base_hi"] = base & (2**64 - 1), base >> 64
    c["lim"], c["fmax1"], c["mod1"], c["nres1"] = lim, fmax1, mod1, len(res)
    c["res"][0, :len(res)] = res
    return c


def tri_sets():
    t504 = np.zeros(8, np.uint64)
    t247 = np.zeros(4, np.uint64)
    a, b = C.c_int(), C.c_int()
    _lib.or_build_tri(t504.ctypes.data_as(_P), t247.ctypes.data_as(_P), C.byref(a), C.byref(b))
    return t504, t247, a.value, b.value


def twin_run(cands, table, use_gate=True, use_tri=True, c4_div=3, threads=8, cap=1 << 20, tile=0, strict=False):
    """k_find4_cls5 semantics: counters dict and sorted hits (tile > 0: the
    reference launch's unbounded c4 tiles of that width; strict: the strict
    residue gate of this build before every query)"""
    cands = np.ascontiguousarray(cands, dtype=TWIN_CAND_DTYPE)
    c = TwinCounts()
    hits = np.zeros(cap, TWIN_HIT_DTYPE)
    nh = C.c_uint64()
    _lib.or_twin_run(cands.ctypes.data_as(_P), len(cands), C.byref(table.x) if table is not None else None,
                     (1 if use_gate else 0) | (2 if strict else 0), 1 if use_tri else 0, c4_div, C.byref(c),
                     hits.ctypes.data_as(_P),
                     cap, C.byref(nh), threads, tile)
    hits = np.sort(hits[:min(nh.value, cap)], order=["cand", "d", "c4", "c3"])
    return {f: getattr(c, f) for f, _ in TwinCounts._fields_}, hits, nh.value


def twin_verify(cand, d, c4, c3, fp, N, cap=16):
    cand = np.ascontiguousarray(cand, dtype=TWIN_CAND_DTYPE).reshape(1)
    terms = np.zeros(cap * 7, np.int64)
    g = np.zeros(cap, np.int64)
    n = _lib.or_twin_verify(cand.ctypes.data_as(_P), d, c4, c3, fp, N, terms.ctypes.data_as(_P),
                            g.ctypes.data_as(_P), cap)
    return [(terms[7 * i:7 * i + 7].tolist(), int(g[i])) for i in range(min(n, cap))]


# ---------------------------------------------------------------------------
# exact sorted pair table (QuadSumIndex, quad_sum.hpp:8-45), Python integers
# ---------------------------------------------------------------------------
def exact_table(N, extra=()):
    """Every a^6 + b^6 with 1 <= a <= b <= N sorted by sum (QuadSumIndex::build,
    quad_sum.hpp:15-30), in the entry form value * 2^s + a (s = bit length of N;
    a = 0 for an extra value).  Returns (entries, s).  Small N only."""
    s = N.bit_length()
    p = [a ** 6 for a in range(N + 1)]
    ent = [((p[a] + p[b]) << s) | a for a in range(1, N + 1) for b in range(a, N + 1)]
    ent += [int(v) << s for v in extra]
    ent.sort()
    return ent, s


def exact_decode(entry, s):
    """(value, a, b) of one entry; b = 0 for an extra value"""
    v, a = entry >> s, entry & ((1 << s) - 1)
    if not a:
        return v, 0, 0
    r = v - a ** 6
    b = round(r ** (1.0 / 6.0))
    while b ** 6 > r:
        b -= 1
    while (b + 1) ** 6 <= r:
        b += 1
    assert b ** 6 == r
    return v, a, b


def exact_range(entries, s, T):
    """equal_sum_range (quad_sum.hpp:37-45) over the entry form: [lo, hi)"""
    import bisect
    return bisect.bisect_left(entries, T << s), bisect.bisect_left(entries, (T + 1) << s)


_lib.or_strict_pass.argtypes = [C.c_uint32] * 5
_lib.or_strict_pass.restype = C.c_int


def strict_pass_T(T):
    """the strict gate on an exact T (an option of this build, not the reference's)"""
    return bool(_lib.or_strict_pass(T % 64, T % 81, T % (31 * 37 * 43), T % (49 * 61 * 67), T % (73 * 79)))
