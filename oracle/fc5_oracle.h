/*
 * fc5_oracle.h -- CPU restatement of the reference's class-5 free-peel path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: it is
 * linked or loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else.  The product (six-one-five_amd/) never
 * calls it; the product fails loudly when its HIP library is missing.
 *
 * Every function restates the reference algorithm in plain C (gcc, with the
 * unsigned __int128 extension) and cites the reference file:line it follows.
 * Arithmetic is exact integer throughout; 192/256-bit values are little-endian
 * arrays of u64 limbs (limb 0 lowest).
 *
 * Parity pinning (what checks each function against something independent of
 * this file) is listed in the header comment of fc5_oracle.c.
 */
#ifndef FC5_ORACLE_H
#define FC5_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- class contracts: fourcore_classes_v3.hpp ------------------------- */
int64_t or_mod_pow6(int64_t base, int64_t mod);
/* number of Stage-1 seeds of class cls for B (classes:114-135); writes up to cap */
int or_seeds_for_B(int64_t B, int cls, int64_t *out, int cap);
/* Stage-1 units of class cls for B under u-band [lo_frac,hi_frac) (classes:137-154) */
int or_unit_candidates(int64_t B, int cls, double lo_frac, double hi_frac,
                       int64_t *out, int cap);
/* eligible B (classes:80-82) */
int or_admissible_B(int64_t B);
/* cls5 free-term residue specs (classes:211-219): e (mod 729) then d (mod 64) */
int or_free_de_cls5(int64_t B, int64_t u, int64_t *e_res, int *n_e, int64_t *e_mod,
                    int64_t *e_fmax, int64_t *d_res, int *n_d, int64_t *d_mod,
                    int64_t *d_fmax);
/* cls2/3/4 single free term (classes:188-209) */
int or_free_d_cls234(int64_t B, int64_t u, int cls, int64_t *res, int *n,
                     int64_t *mod, int64_t *fmax);
/* Stage-1 emitter over [lo,hi] for the classes in cls_mask (bit c = class c);
 * writes (cls,B,u) triples, returns the count (huntv3:371-381).  If out is
 * NULL only counts. */
int64_t or_emit_units(int64_t lo, int64_t hi, unsigned cls_mask, double lo_frac,
                      double hi_frac, int64_t *out_cls, int64_t *out_B,
                      int64_t *out_u, int64_t cap);

/* ---- 192-bit helpers: fourcore_cls5_gpu_v4.cu:77-108,270-308 ----------- */
void or_pow6_192(uint32_t x, uint64_t o[3]);
int or_sub192(uint64_t a[3], const uint64_t b[3]); /* returns borrow */
uint64_t or_mod192(const uint64_t a[3], uint64_t m, uint64_t two64_mod_m);

/* ---- gate: fourcore_cls5_gpu_v4.cu:111-137 ----------------------------- */
typedef struct {
  uint64_t g504[8];
  uint64_t g247[4];
  uint16_t p6504[504];
  uint16_t p6247[247];
} or_gate_t;
void or_build_gate(or_gate_t *g);

/* ---- mixer / packing / query: mix64.hpp:13-25, xor_pack.hpp, xor_filter.hpp */
uint64_t or_mix64(uint64_t x);
uint64_t or_load_cell(const uint8_t *p, uint64_t cell, uint32_t r);
void or_store_cell(uint8_t *p, uint64_t cell, uint32_t r, uint64_t value);
size_t or_packed_bytes(uint64_t m_cells, uint32_t r);
size_t or_capacity_for(size_t n);

/* StoreHeader v2 (store_header.hpp:23-37): 64 bytes, little-endian */
typedef struct {
  uint32_t magic, version, kind, r, w, N;
  uint64_t n_keys, m_cells, mix_seed;
  uint32_t shard_count, shard_index, shard_mode, reserved;
} or_header_t;

typedef struct {
  or_header_t hdr;
  uint8_t *packed;
  size_t packed_len;
} or_xor_t;

int or_xor_might_contain(const or_xor_t *f, uint64_t key);
/* byte-identical builder (xor_filter.hpp:256-470).  keys are copied, sorted
 * and deduplicated.  Returns 0 on success, -1 if every seed stalls. */
int or_build_xor(const uint64_t *keys, size_t n, uint32_t r, uint64_t base_seed,
                 int max_seed_tries, or_xor_t *out);
/* pair table builder with the outer seed schedule (fourcore_xor_store.hpp:67-119);
 * extra keys (planted keys) are appended to the pair keys when n_extra > 0. */
int or_build_pairs(int N, uint32_t r, const uint64_t *extra, size_t n_extra,
                   or_xor_t *out);
void or_xor_free(or_xor_t *f);
int or_xor_save(const char *path, const or_xor_t *f);
/* returns 0 ok, -1 cannot open, -2 bad header/kind, -3 version, -4 N mismatch,
 * -5 short read (fourcore_xor_store.hpp:147-181) */
int or_xor_load(const char *path, or_xor_t *f, int expected_N);
int or_fpr_smoke(const or_xor_t *f, int trials);

/* ---- units: fourcore_cls5_gpu_v4.cu:143-184,346-351 --------------------- */
typedef struct {
  uint64_t B, u;
  uint64_t base[3];
  uint32_t lim, mod1, mod2, fmax1, fmax2, nres1, nres2;
  uint32_t res[24];
} or_unit5_t;
int or_pack_unit5(uint64_t B, uint64_t u, or_unit5_t *U);

/* ---- the strict gate (no counterpart in the reference): 1 when T mod 64,
 * 81, 31, 37, 43, 49, 61, 67, 73 and 79 are each a sum of two sixth powers
 * modulo that number; t64 = T mod 64, t81 = T mod 81, tA = T mod 31*37*43,
 * tB = T mod 49*61*67, tC = T mod 73*79 ---- */
int or_strict_pass(uint32_t t64, uint32_t t81, uint32_t tA, uint32_t tB, uint32_t tC);

/* ---- the peel: k_cls5_192 semantics (fourcore_cls5_gpu_v4.cu:353-410) --- */
typedef struct {
  uint64_t calls, neg, zero, gated, probes, maybe, side_maybe;
  uint64_t queries; /* use_gate & 2 (strict gate): probes that pass it and are queried */
} or_counts_t;

typedef struct {
  uint32_t unit, d, e, flags; /* flags bit0: main filter, bit1: side filter */
  uint64_t fp;
} or_hit_t;

/* Runs the peel for units[0..n).  use_gate bit 0: the reference gate; bit 1:
 * the strict gate of this build (T mod 64, 81, 31, 37, 43, 49, 61, 67, 73 and 79
 * a sum of two sixth powers) on every probe before its query.  side may be NULL (planted-key side filter,
 * probed with the same fingerprint as the main filter).  Hits are written to
 * hits (up to cap; *n_hits gets the true count).  OpenMP over units when
 * threads > 1.  Returns 0. */
int or_cls5_peel(const or_unit5_t *units, int n, const or_xor_t *f,
                 const or_xor_t *side, int use_gate, or_counts_t *counts,
                 or_hit_t *hits, uint64_t cap, uint64_t *n_hits, int threads);

/* ---- classes 2-4 find3 leaf (fourcore_find_v4.cu:165-205,255-284,401-425) ---- */
typedef struct {
  uint64_t jobs, failT, calls, gated, probes, maybe;
  uint64_t queries; /* use_gate & 2: probes that pass the strict gate */
} or_f3counts_t;
/* expand each (cls,B,u) unit into jobs T = (B^6-u^6-(f d)^6)/42^6 (failT when
 * not positive/divisible), window [lo,hi] per job, c3 sweep with the gate on
 * T - c3^6 and probe key (T - c3^6) mod 2^64.  Hits: unit index, d, c3 (in the
 * e field), fp. */
int or_find3(const int *cls, const uint64_t *B, const uint64_t *u, int n, const or_xor_t *f,
             int use_gate, or_f3counts_t *counts, or_hit_t *hits, uint64_t cap, uint64_t *n_hits,
             int threads);
/* exact T = (B^6-u^6-(f d)^6)/42^6 (fourcore_gmp.hpp:45-52) */
int or_T_peel1(uint64_t B, uint64_t u, uint64_t f, uint64_t d, uint64_t *T_lo, uint64_t *T_hi);
/* find3 host verify (fourcore_find_v4.cu:547-551 + report_solution :427-457) */
int or_verify_find3(int cls, uint64_t B, uint64_t u, uint32_t d, uint32_t c, uint64_t fp, int N,
                    int64_t *terms_out, int64_t *gcd_out, int cap);

/* exact T for a grid point (fourcore_gmp.hpp:55-62); returns 1 and T (lo,hi)
 * if num>0, 42^6 | num and T < 2^128, else 0 */
int or_T_peel2(uint64_t B, uint64_t u, uint64_t d, uint64_t e, uint64_t *T_lo,
               uint64_t *T_hi);
/* host verify of one raw hit (fourcore_cls5_gpu_v4.cu:513-542): writes up to
 * cap solutions as 5 sorted terms + gcd; returns count */
int or_verify_hit(uint64_t B, uint64_t u, uint32_t lim, uint32_t d, uint32_t e,
                  uint64_t fp, int N, int64_t *terms_out, int64_t *gcd_out,
                  int cap);

/* ---- (6,1,7) Branch-A class-5 twin: solve_617_v1.cu ----------------------
 * B^6 = u^6 + w^6 + (21d)^6 + 42^6 (c1^6 + c2^6 + c3^6 + c4^6) with
 * u = 14u' (the term not divisible by 3), w = 6v' (the term not divisible by 7)
 * and the grid term 21d (the odd term), d over residues mod 64. */
typedef struct {
  uint64_t B, u, w;
  uint64_t base_lo, base_hi; /* B^6 - u^6 - w^6 */
  uint32_t lim, fmax1, mod1, nres1;
  uint32_t res[8];
} or_twin_cand_t;
/* gen_class_cands cls5 (solve_617_v1.cu:885-906): candidates of B in the u band,
 * u' ascending then v' ascending; writes up to cap, returns the total count */
int or_twin_cands(int64_t B, double u_lo, double u_hi, or_twin_cand_t *out, int cap);
/* triple-sum sets mod 504 and 247 (solve_617_v1.cu:334-346) */
void or_build_tri(uint64_t t504[8], uint64_t t247[4], int *n504, int *n247);
typedef struct {
  uint64_t jobs;            /* (cand, d) cells with a non-empty c4 window */
  uint64_t c4_total;        /* c4 values that reached the triple-sum gate */
  uint64_t c4_tri_skipped;  /* c4 values the triple-sum gate removed */
  uint64_t calls;           /* probe() invocations (pair gate passed) */
  uint64_t gated;           /* c3 values the pair gate removed */
  uint64_t probes;          /* table cell reads, 3 per call */
  uint64_t maybe;           /* maybe-hits of the table */
  uint64_t queries;         /* use_gate & 2 (strict gate): calls that pass it and are queried */
} or_twin_counts_t;
typedef struct { uint32_t cand, d, c4, c3; uint64_t fp; } or_twin_hit_t;
/* k_find4_cls5 semantics (solve_617_v1.cu:563-657).  c4_div: 3 = the reference
 * kernel's c4 lower bound (3 c4^6 >= T), 4 = the exact bound for the largest of
 * four cores (4 c4^6 >= T).  tile = 0 visits every (cand, d, c4) once; tile > 0
 * restates the reference launch's c4 tiles of that width (2048 in the
 * reference): block z starts at lo4 + z * tile with no tile end, so a c4 at
 * offset k from lo4 is visited min(zmax, k / tile + 1) times. */
int or_twin_run(const or_twin_cand_t *cands, int n, const or_xor_t *f, int use_gate, int use_tri,
                int c4_div, or_twin_counts_t *counts, or_twin_hit_t *hits, uint64_t cap,
                uint64_t *n_hits, int threads, int tile);
/* expand_xor_hits + verify_hit cls5 (solve_617_v1.cu:1019-1110): writes up to cap
 * solutions as 7 sorted terms + gcd; returns the count (before deduplication) */
int or_twin_verify(const or_twin_cand_t *C, uint32_t d, uint32_t c4, uint32_t c3, uint64_t fp,
                   int N, int64_t *terms_out, int64_t *gcd_out, int cap);

/* inverse of 21^6 mod 2^64 (fourcore_cls5_gpu_v4.cu:62-66) */
uint64_t or_inv216(void);
/* splitmix64 step (public-domain generator by S. Vigna) used for plants */
uint64_t or_splitmix64(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
