"""Python binding of libfc5.so (include/fc5.h) over ctypes.

This mirrors the reference's in-process seam ``run_cls5_gpu``
(fourcore_cls5_gpu_v4.cu:412-413) and its CLI surface: load or build the
pair table, pack Stage-1 units into a band, run the class-5 free peel on the
GPU, read counters, hits and verified solutions.  It is a thin wrapper: every
computation happens in libfc5.so (HIP kernels for gfx950); when the library
is missing this module raises at import time instead of falling back to
anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libfc5.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libfc5.so not built: {LIB_PATH} (run __graft_entry__.build())")

_lib = C.CDLL(LIB_PATH)

OK, E_ARG, E_IO, E_FORMAT, E_NMISMATCH, E_FPR, E_HIP, E_NOMEM, E_STATE, E_OVERFLOW, E_BUILD, E_PENDING = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11)
MODE_PLAIN, MODE_BATCHED, MODE_CHUNKED = 0, 1, 2
NOALLOC_UNITS, NOALLOC_HITS, NOALLOC_BMAX = 65536, 65536, 5505024  # limits of fc5_run_units_noalloc (fc5.h)


class StoreHeader(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("magic", "version", "kind", "r", "w", "N")] + \
               [(n, C.c_uint64) for n in ("n_keys", "m_cells", "mix_seed")] + \
               [(n, C.c_uint32) for n in ("shard_count", "shard_index", "shard_mode", "reserved")]


class Unit(C.Structure):
    _fields_ = [("B", C.c_uint64), ("u", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "calls", "neg", "zero", "gated", "probes", "maybe", "side_maybe", "n_hits",
        "hits_dropped", "solutions", "units", "units_skipped", "units_irregular", "jobs", "fail_t")] + \
               [("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("strict_out", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tune(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("rows_per_item", "blocks_per_cu", "n_tried", "pad")] + \
               [("best_ms", C.c_double), ("default_ms", C.c_double)]


class Hit(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("unit", "d", "e", "flags")] + [("fp", C.c_uint64)]


class Solution(C.Structure):
    _fields_ = [("B", C.c_uint64), ("u", C.c_uint64)] + \
               [(n, C.c_uint32) for n in ("cls", "d", "e", "a", "b", "c")] + \
               [("terms", C.c_int64 * 5), ("gcd", C.c_int64)]


class TwinResult(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "jobs", "c4_total", "c4_tri_skipped", "calls", "gated", "probes", "maybe", "n_hits", "hits_dropped",
        "solutions", "cands", "items")] + [("kernel_ms", C.c_double), ("wall_ms", C.c_double),
                                           ("strict_out", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TwinSolution(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("B", "u", "w")] + \
               [(n, C.c_uint32) for n in ("d", "c4", "c3", "a", "b", "pad")] + \
               [("terms", C.c_int64 * 7), ("gcd", C.c_int64)]


class ServeOpts(C.Structure):
    _fields_ = [("max_bands", C.c_int), ("rows_per_item", C.c_int), ("arena_units5", C.c_uint64),
                ("arena_units3", C.c_uint64), ("round_points", C.c_double), ("mem_budget", C.c_uint64)]


class ServeTiming(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double), ("rounds", C.c_uint64),
                ("t_release", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "admitted", "done", "pending", "active", "rounds",
                                          "arena5_used", "arena3_used", "device_bytes")] + [("busy_ms", C.c_double)]


# fc5_twin_cand / fc5_twin_hit as numpy records (88 and 24 bytes)
TWIN_CAND_DTYPE = np.dtype([("B", "<u8"), ("u", "<u8"), ("w", "<u8"), ("base_lo", "<u8"), ("base_hi", "<u8"),
                            ("lim", "<u4"), ("fmax1", "<u4"), ("mod1", "<u4"), ("nres1", "<u4"),
                            ("res", "<u4", (8,))])
TWIN_HIT_DTYPE = np.dtype([("cand", "<u4"), ("d", "<u4"), ("c4", "<u4"), ("c3", "<u4"), ("fp", "<u8")])

_P = C.c_void_p
_sig = {
    "fc5_abi_version": (C.c_int, []),
    "fc5_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "fc5_ctx_destroy": (None, [_P]),
    "fc5_last_error": (C.c_char_p, [_P]),
    "fc5_set_option": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "fc5_table_load": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "fc5_table_upload": (C.c_int, [_P, C.POINTER(StoreHeader), _P, C.c_size_t]),
    "fc5_table_build": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_size_t, C.c_int]),
    "fc5_table_info": (C.c_int, [_P, C.POINTER(StoreHeader)]),
    "fc5_table_save": (C.c_int, [_P, C.c_char_p]),
    "fc5_table_download": (C.c_int, [_P, C.POINTER(StoreHeader), _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_table_query": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "fc5_table_fpr_smoke": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "fc5_side_build": (C.c_int, [_P, _P, C.c_size_t, C.c_int]),
    "fc5_side_clear": (C.c_int, [_P]),
    "fc5_band_create": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(_P)]),
    "fc5_band_create_mixed": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(_P)]),
    "fc5_band_destroy": (C.c_int, [_P, _P]),
    "fc5_last_class_result": (C.c_int, [_P, C.c_int, C.POINTER(Result)]),
    "fc5_autotune": (C.c_int, [_P, _P, C.c_int, C.POINTER(Tune)]),
    "fc5_band_run": (C.c_int, [_P, _P, C.c_int, C.POINTER(Result)]),
    "fc5_run_units": (C.c_int, [_P, _P, C.c_size_t, C.c_int, C.POINTER(Result)]),
    "fc5_run_units_noalloc": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(Result), _P, C.c_size_t,
                                        C.POINTER(C.c_size_t), _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_last_hits": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_last_solutions": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_units_calls": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_uint64)]),
    "fc5_measure_gathers": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fc5_measure_copy": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double)]),
    "fc5_measure_gathers_region": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    "fc5_measure_read": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double)]),
    "fc5_measure_valu": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "fc5_measure_partitioned": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "fc5_part_stats": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "fc5_emit_units": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint, C.c_double, C.c_double,
                                 _P, _P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_build_pairs_file": (C.c_int, [C.c_int, C.c_int, C.c_char_p, C.POINTER(StoreHeader)]),
    "fc5_serve_start": (C.c_int, [_P, C.POINTER(ServeOpts)]),
    "fc5_serve_submit": (C.c_int, [_P, _P, _P, C.c_size_t, C.POINTER(C.c_uint64)]),
    "fc5_serve_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(Result)]),
    "fc5_serve_wait": (C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(Result)]),
    "fc5_serve_class_result": (C.c_int, [_P, C.c_uint64, C.c_int, C.POINTER(Result)]),
    "fc5_serve_hits": (C.c_int, [_P, C.c_uint64, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_serve_solutions": (C.c_int, [_P, C.c_uint64, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_serve_timing_of": (C.c_int, [_P, C.c_uint64, C.POINTER(ServeTiming)]),
    "fc5_serve_release": (C.c_int, [_P, C.c_uint64]),
    "fc5_serve_stats_get": (C.c_int, [_P, C.POINTER(ServeStats)]),
    "fc5_serve_stop": (C.c_int, [_P]),
    "fc5_pair_search": (C.c_int, [_P, _P, _P, C.c_size_t, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_exact_build": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "fc5_exact_clear": (C.c_int, [_P]),
    "fc5_exact_info": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_uint32)]),
    "fc5_exact_read": (C.c_int, [_P, C.c_uint64, C.c_size_t, _P]),
    "fc5_exact_lookup": (C.c_int, [_P, _P, _P, C.c_size_t, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_emit_units_device": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint, C.c_double, C.c_double, _P, _P, _P,
                                        C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_band_create_range": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint, C.c_double, C.c_double,
                                        C.POINTER(_P)]),
    "fc5_serve_submit_range": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint, C.c_double, C.c_double,
                                         C.POINTER(C.c_uint64)]),
    "fc5_twin_cands": (C.c_int, [C.c_uint64, C.c_double, C.c_double, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_twin_run": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(TwinResult)]),
    "fc5_twin_run_range": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.POINTER(TwinResult)]),
    "fc5_twin_last_hits": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fc5_twin_last_solutions": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
}
EXPORTS = tuple(_sig)
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class Fc5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fc5 error {code}: {msg}")
        self.code = code


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None and len(a) else None


def units_array(B, u):
    """Pack parallel B, u sequences into the fc5_unit array layout."""
    arr = np.zeros(len(B), dtype=np.dtype([("B", "<u8"), ("u", "<u8")]))
    arr["B"] = np.asarray(B, dtype=np.uint64)
    arr["u"] = np.asarray(u, dtype=np.uint64)
    return arr


def emit_units(lo, hi, classes=(5,), u_band=(0.0, 1.0)):
    """Stage-1 units (cls, B, u) for eligible B in [lo, hi] (fourcore_hunt_v3.cpp:371-381)."""
    mask = 0
    for c in classes:
        mask |= 1 << c
    n = C.c_size_t()
    rc = _lib.fc5_emit_units(lo, hi, mask, u_band[0], u_band[1], None, None, None, 0, C.byref(n))
    if rc:
        raise Fc5Error(rc, "emit_units")
    cls = np.zeros(n.value, np.uint32)
    Bs = np.zeros(n.value, np.uint64)
    us = np.zeros(n.value, np.uint64)
    rc = _lib.fc5_emit_units(lo, hi, mask, u_band[0], u_band[1], _ptr(cls), _ptr(Bs), _ptr(us),
                             n.value, C.byref(n))
    if rc:
        raise Fc5Error(rc, "emit_units")
    return cls, Bs, us


def units_calls(units):
    out = C.c_uint64()
    rc = _lib.fc5_units_calls(_ptr(units), len(units), C.byref(out))
    if rc:
        raise Fc5Error(rc, "units_calls")
    return out.value


def twin_cands(B, u_band=(0.0, 1.0)):
    """(6,1,7) class-5 candidates of one B (gen_class_cands, solve_617_v1.cu:885-906)."""
    n = C.c_size_t()
    rc = _lib.fc5_twin_cands(B, u_band[0], u_band[1], None, 0, C.byref(n))
    if rc:
        raise Fc5Error(rc, "fc5_twin_cands")
    out = np.zeros(n.value, TWIN_CAND_DTYPE)
    if n.value:
        _lib.fc5_twin_cands(B, u_band[0], u_band[1], out.ctypes.data_as(_P), n.value, C.byref(n))
    return out


def build_pairs_file(N, r, path):
    h = StoreHeader()
    rc = _lib.fc5_build_pairs_file(N, r, path.encode(), C.byref(h))
    if rc:
        raise Fc5Error(rc, "build_pairs_file")
    return h


@dataclass
class RunOutput:
    result: dict
    hits: np.ndarray
    solutions: list
    rc: int


HIT_DTYPE = np.dtype([("unit", "<u4"), ("d", "<u4"), ("e", "<u4"), ("flags", "<u4"), ("fp", "<u8")])
HIT_EXTRA = 4  # hit flag: the exact check found the target among the exact table's extra values


class Band:
    """units: fc5_unit array; cls: optional per-unit class (2..5), default all 5"""

    def __init__(self, ctx, units, cls=None):
        self.ctx = ctx
        self.units = np.ascontiguousarray(units)
        self.h = _P()
        if cls is None:
            rc = _lib.fc5_band_create(ctx.h, _ptr(self.units), len(self.units), C.byref(self.h))
        else:
            self.cls = np.ascontiguousarray(cls, dtype=np.uint32)
            rc = _lib.fc5_band_create_mixed(ctx.h, _ptr(self.cls), _ptr(self.units), len(self.units),
                                            C.byref(self.h))
        ctx._check(rc)

    @classmethod
    def from_range(cls_, ctx, lo, hi, classes=(5,), u_band=(0.0, 1.0)):
        """a band straight from a B range (Stage-1 on the device)"""
        b = cls_.__new__(cls_)
        b.ctx = ctx
        b.h = _P()
        mask = sum(1 << c for c in classes)
        ctx._check(_lib.fc5_band_create_range(ctx.h, lo, hi, mask, u_band[0], u_band[1], C.byref(b.h)))
        return b

    def run(self, mode=MODE_BATCHED, allow_overflow=False):
        res = Result()
        rc = _lib.fc5_band_run(self.ctx.h, self.h, mode, C.byref(res))
        if rc not in (OK, E_OVERFLOW) or (rc == E_OVERFLOW and not allow_overflow):
            self.ctx._check(rc)
        out = RunOutput(res.as_dict(), self.ctx.last_hits(), self.ctx.last_solutions(), rc)
        out.by_class = {k: self.ctx.class_result(k) for k in (2, 3, 4, 5)}
        return out

    def autotune(self, sample_every=16):
        t = Tune()
        self.ctx._check(_lib.fc5_autotune(self.ctx.h, self.h, sample_every, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tune._fields_ if f != "pad"}

    def close(self):
        if self.h:
            _lib.fc5_band_destroy(self.ctx.h, self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device context (fc5_ctx): tables, options, runs."""

    def __init__(self, device=0, **options):
        self.h = _P()
        rc = _lib.fc5_ctx_create(device, C.byref(self.h))
        if rc:
            raise Fc5Error(rc, f"cannot create a context on device {device}")
        for k, v in options.items():
            self.set_option(k, v)

    def _check(self, rc):
        if rc != OK:
            raise Fc5Error(rc, _lib.fc5_last_error(self.h).decode())

    def set_option(self, key, value):
        self._check(_lib.fc5_set_option(self.h, key.encode(), int(value)))

    def table_load(self, path, expected_N=0):
        self._check(_lib.fc5_table_load(self.h, path.encode(), expected_N))

    def table_upload(self, header: StoreHeader, packed: bytes):
        buf = np.frombuffer(packed, dtype=np.uint8)
        self._check(_lib.fc5_table_upload(self.h, C.byref(header), _ptr(buf), len(buf)))

    def table_build(self, N, r=48, extra=None, on_gpu=False):
        ex = np.ascontiguousarray(extra if extra is not None else np.zeros(0, np.uint64), dtype=np.uint64)
        self._check(_lib.fc5_table_build(self.h, N, r, _ptr(ex), len(ex), 1 if on_gpu else 0))

    def table_info(self):
        h = StoreHeader()
        self._check(_lib.fc5_table_info(self.h, C.byref(h)))
        return h

    def table_save(self, path):
        self._check(_lib.fc5_table_save(self.h, path.encode()))

    def table_download(self):
        """(header, packed bytes as a numpy uint8 array) in the file layout"""
        h = StoreHeader()
        n = C.c_size_t()
        self._check(_lib.fc5_table_download(self.h, C.byref(h), None, 0, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        self._check(_lib.fc5_table_download(self.h, C.byref(h), _ptr(buf), n.value, C.byref(n)))
        return h, buf

    def table_query(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(len(keys), np.uint8)
        self._check(_lib.fc5_table_query(self.h, _ptr(keys), len(keys), _ptr(out)))
        return out

    def fpr_smoke(self, trials=200000):
        h = C.c_int()
        self._check(_lib.fc5_table_fpr_smoke(self.h, trials, C.byref(h)))
        return h.value

    def side_build(self, keys, r=48):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self._check(_lib.fc5_side_build(self.h, _ptr(keys), len(keys), r))

    def side_clear(self):
        self._check(_lib.fc5_side_clear(self.h))

    def measure_gathers(self, n):
        """(ms for n full queries, ms for 3n raw random cell loads) on the main table"""
        a, b = C.c_double(), C.c_double()
        self._check(_lib.fc5_measure_gathers(self.h, n, C.byref(a), C.byref(b)))
        return a.value, b.value

    def measure_copy(self, nbytes):
        t = C.c_double()
        self._check(_lib.fc5_measure_copy(self.h, nbytes, C.byref(t)))
        return t.value

    def measure_gathers_region(self, region_cells, n):
        """ms for n raw random cell loads within the first region_cells cells of the main table"""
        t = C.c_double()
        self._check(_lib.fc5_measure_gathers_region(self.h, region_cells, n, C.byref(t)))
        return t.value

    def measure_read(self, nbytes):
        """ms for one read-only streaming pass over a fresh device buffer of nbytes"""
        t = C.c_double()
        self._check(_lib.fc5_measure_read(self.h, nbytes, C.byref(t)))
        return t.value

    def measure_valu(self, kind):
        """integer VALU issue rate (wavefront instructions/s) of microbenchmark kind 0..3"""
        t = C.c_double()
        self._check(_lib.fc5_measure_valu(self.h, kind, C.byref(t)))
        return t.value

    def measure_partitioned(self, n):
        """partitioned probing of n pseudo-random keys: (stage ms, counts) as
        named in fc5_measure_partitioned"""
        ms = (C.c_double * 9)()
        cnt = (C.c_uint64 * 8)()
        self._check(_lib.fc5_measure_partitioned(self.h, n, ms, cnt))
        names = ("first", "fine0", "lookup0", "fine1", "lookup1", "fine2", "last", "check", "direct")
        cn = ("accepted", "accepted_hsum", "direct_accepted", "direct_hsum", "full_queries", "lost", "slices",
              "coarse")
        return dict(zip(names, list(ms))), dict(zip(cn, list(cnt)))

    def part_stats(self):
        """the last partitioned batched run: host ms of the emit launches, the
        partition pipelines and the identify launch, event ms of the pipeline
        stages summed over the pieces; pieces, accepted hashes, full queries,
        emit chunks filled"""
        ms = (C.c_double * 11)()
        cnt = (C.c_uint64 * 4)()
        self._check(_lib.fc5_part_stats(self.h, ms, cnt))
        return (dict(zip(("emit", "pipeline", "identify", "first", "fine0", "lookup0", "fine1", "lookup1", "fine2",
                          "last", "check"), list(ms))),
                dict(zip(("pieces", "accepted", "full_queries", "chunks"), list(cnt))))

    def band(self, units, cls=None):
        return Band(self, units, cls)

    def class_result(self, cls):
        r = Result()
        self._check(_lib.fc5_last_class_result(self.h, cls, C.byref(r)))
        return r.as_dict()

    def run(self, units, mode=MODE_BATCHED, allow_overflow=False, cls=None):
        b = Band(self, units, cls)
        try:
            return b.run(mode, allow_overflow)
        finally:
            b.close()

    def last_hits(self):
        n = C.c_size_t()
        self._check(_lib.fc5_last_hits(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, HIT_DTYPE)
        if n.value:
            self._check(_lib.fc5_last_hits(self.h, out.ctypes.data_as(_P), n.value, C.byref(n)))
        return out

    def run_noalloc(self, units, hit_cap=1 << 16, sol_cap=64, allow_overflow=False):
        """fc5_run_units_noalloc: the batched class-5 search with no device allocation
        (RunOutput; hits beyond hit_cap are not copied out, their count is in n_hits)"""
        units = np.ascontiguousarray(units)
        res = Result()
        hits = np.zeros(hit_cap, HIT_DTYPE)
        sols = (Solution * max(1, sol_cap))()
        nh, ns = C.c_size_t(), C.c_size_t()
        rc = _lib.fc5_run_units_noalloc(self.h, _ptr(units), len(units), C.byref(res), _ptr(hits), hit_cap,
                                        C.byref(nh), C.cast(sols, _P), sol_cap, C.byref(ns))
        if rc not in (OK, E_OVERFLOW) or (rc == E_OVERFLOW and not allow_overflow):
            self._check(rc)
        sl = [dict(B=x.B, u=x.u, cls=x.cls, d=x.d, e=x.e, a=x.a, b=x.b, c=x.c, terms=list(x.terms), gcd=x.gcd)
              for x in sols[:min(ns.value, sol_cap)]]
        return RunOutput(res.as_dict(), hits[:min(nh.value, hit_cap)], sl, rc)

    def last_solutions(self):
        n = C.c_size_t()
        self._check(_lib.fc5_last_solutions(self.h, None, 0, C.byref(n)))
        arr = (Solution * max(1, n.value))()
        if n.value:
            self._check(_lib.fc5_last_solutions(self.h, C.cast(arr, _P), n.value, C.byref(n)))
        return [dict(B=s.B, u=s.u, cls=s.cls, d=s.d, e=s.e, a=s.a, b=s.b, c=s.c, terms=list(s.terms),
                     gcd=s.gcd) for s in arr[:n.value]]

    def pair_search(self, targets, hi):
        """device exact check: list of (i, a, b) with a <= b <= hi[i], a^6 + b^6 = targets[i]"""
        t = np.zeros(2 * len(targets), np.uint64)
        for i, v in enumerate(targets):
            t[2 * i], t[2 * i + 1] = v & (2 ** 64 - 1), v >> 64
        h = np.ascontiguousarray(hi, dtype=np.uint32)
        n = C.c_size_t()
        self._check(_lib.fc5_pair_search(self.h, _ptr(t), _ptr(h), len(targets), None, 0, C.byref(n)))
        out = np.zeros(3 * max(1, n.value), np.uint32)
        self._check(_lib.fc5_pair_search(self.h, _ptr(t), _ptr(h), len(targets), out.ctypes.data_as(_P), n.value,
                                         C.byref(n)))
        return [tuple(int(x) for x in out[3 * k:3 * k + 3]) for k in range(n.value)]

    # ---- exact sorted pair table ----
    @staticmethod
    def _words(values):
        t = np.zeros(2 * max(1, len(values)), np.uint64)
        for i, v in enumerate(values):
            t[2 * i], t[2 * i + 1] = int(v) & _M64, int(v) >> 64
        return t

    def exact_build(self, N, extra=()):
        """every a^6 + b^6 (1 <= a <= b <= N) and the extra 128-bit values, sorted on the device"""
        t = self._words(list(extra))
        self._check(_lib.fc5_exact_build(self.h, N, _ptr(t) if len(extra) else None, len(extra)))

    def exact_clear(self):
        self._check(_lib.fc5_exact_clear(self.h))

    def exact_info(self):
        """(entries, N, shift s): entry = value * 2^s + a, a = 0 for an extra value"""
        n, N, s = C.c_uint64(), C.c_int(), C.c_uint32()
        self._check(_lib.fc5_exact_info(self.h, C.byref(n), C.byref(N), C.byref(s)))
        return n.value, N.value, s.value

    def exact_read(self, first, n):
        """entries [first, first + n) as Python ints"""
        out = np.zeros(2 * max(1, n), np.uint64)
        self._check(_lib.fc5_exact_read(self.h, first, n, out.ctypes.data_as(_P)))
        return [int(out[2 * i]) | (int(out[2 * i + 1]) << 64) for i in range(n)]

    def exact_lookup(self, targets, hi):
        """equal ranges: list of (i, a, b) with b <= hi[i], (i, 0, 0) for an extra value"""
        t = self._words(list(targets))
        h = np.ascontiguousarray(hi, dtype=np.uint32)
        n = C.c_size_t()
        self._check(_lib.fc5_exact_lookup(self.h, _ptr(t), _ptr(h), len(targets), None, 0, C.byref(n)))
        out = np.zeros(3 * max(1, n.value), np.uint32)
        self._check(_lib.fc5_exact_lookup(self.h, _ptr(t), _ptr(h), len(targets), out.ctypes.data_as(_P), n.value,
                                          C.byref(n)))
        return [tuple(int(x) for x in out[3 * k:3 * k + 3]) for k in range(n.value)]

    # ---- resident serving loop ----
    def serve_start(self, **opts):
        o = ServeOpts()
        for k, v in opts.items():
            setattr(o, k, v)
        self._check(_lib.fc5_serve_start(self.h, C.byref(o)))

    def serve_submit(self, units, cls=None):
        units = np.ascontiguousarray(units)
        cl = None if cls is None else np.ascontiguousarray(cls, dtype=np.uint32)
        bid = C.c_uint64()
        self._check(_lib.fc5_serve_submit(self.h, _ptr(cl), _ptr(units), len(units), C.byref(bid)))
        return bid.value

    def serve_submit_range(self, lo, hi, classes=(4, 5), u_band=(0.0, 1.0)):
        mask = sum(1 << c for c in classes)
        bid = C.c_uint64()
        self._check(_lib.fc5_serve_submit_range(self.h, lo, hi, mask, u_band[0], u_band[1], C.byref(bid)))
        return bid.value

    def emit_units_device(self, lo, hi, classes=(5,), u_band=(0.0, 1.0)):
        """Stage-1 on the device: (cls, B, u) arrays, as emit_units"""
        mask = sum(1 << c for c in classes)
        n = C.c_size_t()
        self._check(_lib.fc5_emit_units_device(self.h, lo, hi, mask, u_band[0], u_band[1], None, None, None, 0,
                                               C.byref(n)))
        cl = np.zeros(n.value, np.uint32)
        Bs = np.zeros(n.value, np.uint64)
        us = np.zeros(n.value, np.uint64)
        if n.value:
            self._check(_lib.fc5_emit_units_device(self.h, lo, hi, mask, u_band[0], u_band[1], _ptr(cl), _ptr(Bs),
                                                   _ptr(us), n.value, C.byref(n)))
        return cl, Bs, us

    def serve_poll(self, band_id):
        """True when the band is finished"""
        rc = _lib.fc5_serve_poll(self.h, band_id, None)
        if rc == E_PENDING:
            return False
        if rc not in (OK, E_OVERFLOW):
            self._check(rc)
        return True

    def serve_wait(self, band_id, timeout=-1.0):
        rc = _lib.fc5_serve_wait(self.h, band_id, timeout, None)
        if rc == E_PENDING:
            return False
        if rc not in (OK, E_OVERFLOW):
            self._check(rc)
        return True

    def serve_result(self, band_id, allow_overflow=False):
        """RunOutput of a finished band (result, hits, solutions, by_class)"""
        res = Result()
        rc = _lib.fc5_serve_poll(self.h, band_id, C.byref(res))
        if rc not in (OK, E_OVERFLOW) or (rc == E_OVERFLOW and not allow_overflow):
            self._check(rc)
        n = C.c_size_t()
        self._check(_lib.fc5_serve_hits(self.h, band_id, None, 0, C.byref(n)))
        hits = np.zeros(n.value, HIT_DTYPE)
        if n.value:
            self._check(_lib.fc5_serve_hits(self.h, band_id, hits.ctypes.data_as(_P), n.value, C.byref(n)))
        self._check(_lib.fc5_serve_solutions(self.h, band_id, None, 0, C.byref(n)))
        arr = (Solution * max(1, n.value))()
        if n.value:
            self._check(_lib.fc5_serve_solutions(self.h, band_id, C.cast(arr, _P), n.value, C.byref(n)))
        sols = [dict(B=x.B, u=x.u, cls=x.cls, d=x.d, e=x.e, a=x.a, b=x.b, c=x.c, terms=list(x.terms), gcd=x.gcd)
                for x in arr[:n.value]]
        out = RunOutput(res.as_dict(), hits, sols, rc)
        by = {}
        for k in (2, 3, 4, 5):
            r = Result()
            self._check(_lib.fc5_serve_class_result(self.h, band_id, k, C.byref(r)))
            by[k] = r.as_dict()
        out.by_class = by
        return out

    def serve_timing(self, band_id):
        t = ServeTiming()
        rc = _lib.fc5_serve_timing_of(self.h, band_id, C.byref(t))
        if rc not in (OK, E_PENDING):
            self._check(rc)
        return {f: getattr(t, f) for f, _ in ServeTiming._fields_}

    def serve_release(self, band_id):
        self._check(_lib.fc5_serve_release(self.h, band_id))

    def serve_stats(self):
        st = ServeStats()
        self._check(_lib.fc5_serve_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServeStats._fields_}

    def serve_stop(self):
        self._check(_lib.fc5_serve_stop(self.h))

    def twin_run(self, cands, allow_overflow=False):
        """(6,1,7) class-5 twin over candidate records (TWIN_CAND_DTYPE)."""
        cands = np.ascontiguousarray(cands, dtype=TWIN_CAND_DTYPE)
        res = TwinResult()
        rc = _lib.fc5_twin_run(self.h, _ptr(cands), len(cands), C.byref(res))
        return self._twin_out(rc, res, allow_overflow)

    def twin_run_range(self, B_lo, B_hi, u_band=(0.0, 1.0), allow_overflow=False):
        res = TwinResult()
        rc = _lib.fc5_twin_run_range(self.h, B_lo, B_hi, u_band[0], u_band[1], C.byref(res))
        return self._twin_out(rc, res, allow_overflow)

    def _twin_out(self, rc, res, allow_overflow):
        if rc not in (OK, E_OVERFLOW) or (rc == E_OVERFLOW and not allow_overflow):
            self._check(rc)
        n = C.c_size_t()
        self._check(_lib.fc5_twin_last_hits(self.h, None, 0, C.byref(n)))
        hits = np.zeros(n.value, TWIN_HIT_DTYPE)
        if n.value:
            self._check(_lib.fc5_twin_last_hits(self.h, hits.ctypes.data_as(_P), n.value, C.byref(n)))
        self._check(_lib.fc5_twin_last_solutions(self.h, None, 0, C.byref(n)))
        arr = (TwinSolution * max(1, n.value))()
        if n.value:
            self._check(_lib.fc5_twin_last_solutions(self.h, C.cast(arr, _P), n.value, C.byref(n)))
        sols = [dict(B=x.B, u=x.u, w=x.w, d=x.d, c4=x.c4, c3=x.c3, a=x.a, b=x.b, terms=list(x.terms), gcd=x.gcd)
                for x in arr[:n.value]]
        return res.as_dict(), hits, sols

    def close(self):
        if self.h:
            _lib.fc5_ctx_destroy(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# planted keys (workload definition for the plant-recovery checks)
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(state):
    """one step of splitmix64 (S. Vigna); returns (value, new_state)"""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31), state


def _roots(t, m):
    from math import gcd
    return [a for a in range(1, m) if gcd(a, m) == 1 and pow(a, 6, m) == t % m]


def cls5_free_values(B, u):
    """valid e (mod 729 classes, <= (B-1)/14) and d (mod 64 classes, <= (B-1)/21) of a unit"""
    te = pow(B, 6, 729) * pow(pow(14, 6, 729), -1, 729) % 729
    td = (pow(B, 6, 64) - pow(u, 6, 64)) * pow(pow(21, 6, 64), -1, 64) % 64
    es = sorted(e for r in _roots(te, 729) for e in range(r, (B - 1) // 14 + 1, 729))
    ds = sorted(d for r in _roots(td, 64) for d in range(r, (B - 1) // 21 + 1, 64))
    return es, ds


def _pair_sum_set(m):
    p = [pow(a, 6, m) for a in range(m)]
    return frozenset((x + y) % m for x in p for y in p)


_S504, _S247 = _pair_sum_set(504), _pair_sum_set(247)
# the strict gate's sets (option "strict_gate")
_STRICT = {m: _pair_sum_set(m) for m in (64, 81, 31, 37, 43, 49, 61, 67, 73, 79)}


def strict_gate_ok(T):
    """T passes the strict gate: T mod 64, 81, 31, 37, 43, 49, 61, 67, 73, 79 each a sum
    of two sixth powers"""
    return all(T % m in s for m, s in _STRICT.items())


def plant_points(Bs, us, n, seed, exact=False, strict=False):
    """n distinct planted grid points (unit index, e, d, key) drawn with splitmix64:
    the unit uniformly over the band, then (e, d) uniformly over the unit's
    valid values, redrawn while the residual R is not positive or T = R/42^6
    fails the sixth-power residue gate (a real solution always passes it: T is
    a pair sum), at most 2000 draws per unit; key = T mod 2^64.  exact=True
    appends T itself (unit index, e, d, key, T) for the exact table.  strict=True
    also redraws while T fails the strict gate (a pair sum passes that too), so
    the plants are found with the option strict_gate on as well."""
    s = seed & _M64
    out, seen = [], set()
    while len(out) < n:
        v, s = splitmix64(s)
        i = v % len(Bs)
        B, u = int(Bs[i]), int(us[i])
        es, ds = cls5_free_values(B, u)
        base = B ** 6 - u ** 6
        for _ in range(2000):  # units whose positive region is tiny are passed over
            v, s = splitmix64(s)
            e = es[v % len(es)]
            v, s = splitmix64(s)
            d = ds[v % len(ds)]
            R = base - (21 * d) ** 6 - (14 * e) ** 6
            if R <= 0:
                continue
            T = R // 42 ** 6
            if T % 504 in _S504 and T % 247 in _S247 and (not strict or strict_gate_ok(T)):
                break
        else:
            continue
        if (i, e, d) in seen:
            continue
        seen.add((i, e, d))
        out.append((i, e, d, T & _M64, T) if exact else (i, e, d, T & _M64))
    return out
