/*
 * fc5.h -- C ABI of the MI355X class-5 free-peel engine (libfc5.so).
 *
 * The reference has no library boundary: its hot path sits behind the CLI
 * fourcore_cls5_gpu_v4 (fourcore_cls5_gpu_v4.cu:611-709), the in-process seam
 * run_cls5_gpu (fourcore_cls5_gpu_v4.cu:412-413), the Stage-1 emitter
 * (fourcore_hunt_v3.cpp:371-381) and the table writer (xor_build_save.cpp:21-57).
 * Each entry point below names the reference interface it replaces.  All
 * signatures are plain C: pointers, sizes and POD structs; no C++ or torch
 * types.  Every function returns FC5_OK (0) or a negative FC5_E* code; the
 * message for the last error of a context is fc5_last_error(ctx).  Nothing in
 * the library calls exit().
 *
 * Ownership: the caller owns every array it passes (the library copies what it
 * keeps before returning) and every output buffer.  The library owns device
 * memory (tables, units, hits, counters).  One context per host thread.
 */
#ifndef FC5_H
#define FC5_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC5_ABI_VERSION 3

enum {
  FC5_OK = 0,
  FC5_E_ARG = -1,        /* bad argument */
  FC5_E_IO = -2,         /* cannot open / short read / short write */
  FC5_E_FORMAT = -3,     /* bad header, magic, kind or version */
  FC5_E_NMISMATCH = -4,  /* table N differs from the requested N */
  FC5_E_FPR = -5,        /* FPR smoke above 16/200000: refuse to search */
  FC5_E_HIP = -6,        /* HIP runtime error */
  FC5_E_NOMEM = -7,      /* device or host allocation failed */
  FC5_E_STATE = -8,      /* no table loaded, unknown band, ... */
  FC5_E_OVERFLOW = -9,   /* hit buffer overflowed (hits were dropped) */
  FC5_E_BUILD = -10,     /* xor construction stalled at every seed */
  FC5_E_PENDING = -11    /* serving: the band is not finished yet */
};

/* search modes of fc5_band_run */
enum {
  FC5_MODE_PLAIN = 0,   /* reference shape: 16-unit chunks, one synchronous launch each */
  FC5_MODE_BATCHED = 1, /* one persistent launch fed by a device work queue */
  FC5_MODE_CHUNKED = 2  /* the batched kernel in the reference's launch pattern: one
                           synchronous launch per 16 units (fourcore_cls5_gpu_v4.cu:468-501) */
};

/* StoreHeader v2 (spike/ribbon/store_header.hpp:23-37): 64 bytes, little-endian */
typedef struct {
  uint32_t magic, version, kind, r, w, N;
  uint64_t n_keys, m_cells, mix_seed;
  uint32_t shard_count, shard_index, shard_mode, reserved;
} fc5_store_header;

/* one Stage-1 unit of class 5: a ".buc" line "5 B u" (fourcore_cls5_gpu_v4.cu:155-167) */
typedef struct {
  uint64_t B, u;
} fc5_unit;

/* counters of one run.  Reference names (fourcore_cls5_gpu_v4.cu:382-409):
 * calls = every valid (unit,e,d) grid point; gated = gate rejects;
 * probes = gate survivors (xor queries); maybe = xor maybe-hits (nhit).
 * neg/zero split the remaining points (negative / zero residual).
 * Classes 2-4 (find3, fourcore_find_v4.cu:165-284): jobs = free terms d with a
 * positive exact T, fail_t = free terms whose T is not positive/divisible
 * (failT), calls = c3 values swept over the jobs' windows. */
typedef struct {
  uint64_t calls, neg, zero, gated, probes;
  uint64_t maybe, side_maybe;
  uint64_t n_hits, hits_dropped;
  uint64_t solutions;
  uint64_t units, units_skipped, units_irregular;
  uint64_t jobs, fail_t;
  double kernel_ms, wall_ms;
  uint64_t strict_out;  /* option "strict_gate": gate survivors the strict gate rejected
                           (no xor query made); the queries made are probes - strict_out */
} fc5_result;

/* a filter hit (fourcore_cls5_gpu_v4.cu:140).  unit indexes the band's units
 * as passed in; flags bit0 = main table maybe, bit1 = side (planted) table maybe,
 * bit2 = FC5_HIT_EXTRA (exact check found the target among the exact table's extra values).
 * Class 5: d, e are the free terms.  Classes 2-4: d is the free term and e holds
 * the third core c3 (Hit.c of fourcore_find_v4.cu:99). */
typedef struct {
  uint32_t unit, d, e, flags;
  uint64_t fp;
} fc5_hit;

/* a verified solution line "SOLUTION cls=C B= u= a1..a5 gcd=" (fourcore_cls5_gpu_v4.cu:524-536,
 * fourcore_find_v4.cu:427-457): pair (a,b) for class 5, cores (a,b,c) for classes 2-4 */
typedef struct {
  uint64_t B, u;
  uint32_t cls, d, e, a, b, c;
  int64_t terms[5];
  int64_t gcd;
} fc5_solution;

typedef struct fc5_ctx fc5_ctx;
typedef struct fc5_band fc5_band;

int fc5_abi_version(void);

/* ---- context (replaces main's cudaSetDevice and run_cls5_gpu set-up,
 *      fourcore_cls5_gpu_v4.cu:414-466,673) ---- */
int fc5_ctx_create(int device, fc5_ctx **out);
void fc5_ctx_destroy(fc5_ctx *ctx);
/* text of the context's last error; it stays valid until the calling thread's
 * next fc5_last_error call (safe while a server thread runs) */
const char *fc5_last_error(const fc5_ctx *ctx);
/* options: "hit_cap" (default 1<<20), "use_gate" (1; 0 = --no-gate,
 * fourcore_cls5_gpu_v4.cu:634), "strict_gate" (0; 1 = before each xor query, also
 * require T mod 64, 81, 31, 37, 43, 49, 61, 67, 73 and 79 to be sums of two sixth powers:
 * a necessary condition for T = a^6 + b^6 that the reference does not test, so no
 * solution is lost; counters and hits are the reference's except that the xor
 * filter's false positives among the T it rejects are not reported, and
 * strict_out counts the rejected gate survivors; classes 2-5, every mode, and the
 * (6,1,7) twin's batched kernel on its keys T - c4^6 - c3^6), "sweep" (1: the class-5 rows of the batched kernel sieve 32 d
 * values per LDS word, the reference gate split into its conditions mod 8, 9, 7,
 * 13 and 19; 0: one bitmap test of T mod 124488 per grid point; same counters and
 * hits), "blocks_per_cu" (0 = occupancy limit) for the persistent grid, "rows_per_item" (e- or d-rows per work item: 4, 8, 16, 32, or
 * 0, the default: by band size, the largest of 32, 16, 8 that leaves every wave of
 * the persistent grid at least 8 items, else 4), "use_graph" (0; 1 replays the batched pipeline as a captured hipGraph),
 * "load_policy" (-1..3, default -1: cache policy of the xor cell gathers:
 * -1 by table size (plain loads up to 1 GiB of cells, non-temporal above),
 * 0 default, 1 nontemporal, 2 device-scope, 3 system-scope), "verify" (1: exact
 * check of hits: device pair search + host identity check), "verify_table" (0;
 * 1 answers the exact check from the exact sorted pair table below),
 * "twin_*" (the (6,1,7) twin options below; "twin_plain" 1 runs its
 * reference-shaped kernel), "probe_mode" (-1, the default: the batched search
 * probes the filter by partition for tables of 2^28 cells or more with r >= 24,
 * else with three gathers per probe; 0 gathers; 1 by partition; same counters
 * and hits), "part_records" (emit buffer of the partitioned probe, 4096..2^32-1
 * hashes, default 2^32 - 4096), "part_hbm_limit" (bytes the partitioned probe's
 * buffers may take, 0 = no limit; beyond it, or when they cannot be allocated,
 * probe_mode -1 probes with gathers and probe_mode 1 fails with FC5_E_NOMEM),
 * "stop_first" (0, fourcore_cls5_gpu_v4.cu:635) */
int fc5_set_option(fc5_ctx *ctx, const char *key, int64_t value);

/* ---- pair tables (replace acquire_xor_filter + the H2D upload,
 *      fourcore_cls5_gpu_v4.cu:427-440,568-594) ---- */
/* xor_load_file (fourcore_xor_store.hpp:147-181) with the exact-N check
 * (expected_N <= 0 accepts any N) and the FPR smoke refusal (> 16/200000),
 * then upload.  Returns FC5_E_IO / FC5_E_FORMAT / FC5_E_NMISMATCH / FC5_E_FPR;
 * on any failure the current table stays.  Tables of more packed bytes than the
 * option "stream_load_min" (4 GiB) are streamed from the file to the device in
 * chunks (no host copy) and smoked there with the same keys; smaller ones are
 * read whole and a host copy is kept. */
int fc5_table_load(fc5_ctx *ctx, const char *path, int expected_N);
/* upload a table already in host memory (header + packed bytes as on disk) */
int fc5_table_upload(fc5_ctx *ctx, const fc5_store_header *hdr, const uint8_t *packed, size_t len);
/* build the pair table for 1<=a<=b<=N at r bits (xor_build_pairs,
 * fourcore_xor_store.hpp:67-119) plus n_extra extra keys.  on_gpu = 0 builds
 * on the host, byte-identical to the reference builder; on_gpu = 1 builds on
 * the device (a valid xor filter over the same keys, different bytes). */
int fc5_table_build(fc5_ctx *ctx, int N, int r, const uint64_t *extra, size_t n_extra, int on_gpu);
int fc5_table_info(const fc5_ctx *ctx, fc5_store_header *out);
/* write the current table as StoreHeader v2 + packed bytes (xor_save_file,
 * fourcore_xor_store.hpp:126-143) */
int fc5_table_save(fc5_ctx *ctx, const char *path);
/* copy the table out in the file layout (header + r-bit packed bytes, as
 * written by xor_save_file); *len gets the packed size; cap too small ->
 * FC5_E_ARG with *len set */
int fc5_table_download(fc5_ctx *ctx, fc5_store_header *hdr, uint8_t *packed, size_t cap, size_t *len);
/* query the device table for n keys (xor_might_contain, xor_filter.hpp:472-483) */
int fc5_table_query(fc5_ctx *ctx, const uint64_t *keys, size_t n, uint8_t *out);
/* FPR smoke on the device table (xor_fpr_smoke, fourcore_xor_store.hpp:54-65) */
int fc5_table_fpr_smoke(fc5_ctx *ctx, int trials, int *hits);

/* ---- planted-key side table: a small xor filter probed with the same
 *      fingerprint as the main table; no reference counterpart (the reference
 *      plants by fabricating units, solve_516_v3.cu:1149-1170) ---- */
int fc5_side_build(fc5_ctx *ctx, const uint64_t *keys, size_t n, int r);
int fc5_side_clear(fc5_ctx *ctx);

/* ---- bands of units (replace load_buc5 + pack_unit5 + the chunk loop of
 *      run_cls5_gpu, fourcore_cls5_gpu_v4.cu:155-184,468-554) ---- */
/* pack and upload units; units the reference would not pack are skipped and
 * counted.  The band keeps the caller's unit order for hit records. */
int fc5_band_create(fc5_ctx *ctx, const fc5_unit *units, size_t n, fc5_band **out);
/* a band of units of classes 2-5 (the .buc lines "cls B u"); classes 2-4 run
 * the find3 leaf (fourcore_find_v4.cu:165-284) in the same persistent kernel,
 * other classes are skipped and counted */
int fc5_band_create_mixed(fc5_ctx *ctx, const uint32_t *cls, const fc5_unit *units, size_t n,
                          fc5_band **out);
int fc5_band_destroy(fc5_ctx *ctx, fc5_band *band);
/* run one band: all counters, hits and (if "verify") solutions */
int fc5_band_run(fc5_ctx *ctx, fc5_band *band, int mode, fc5_result *res);
/* autotuner (no reference counterpart; the reference fixes 16 units per launch
 * and 256 threads, fourcore_cls5_gpu_v4.cu:468,496): times the persistent
 * kernel on every sample_every-th unit of the band for rows per work item in
 * {4, 8, 16, 32} and two persistent grid sizes, keeps the fastest in the
 * context options "rows_per_item" and "blocks_per_cu" unless it is within 1% of
 * the untuned shape (then the options are left as they were), and reports it
 * (default_ms: the shape used without tuning -- the context's rows_per_item, or
 * the one chosen by the band's size -- at the occupancy-limit grid) */
typedef struct {
  int rows_per_item, blocks_per_cu, n_tried, pad;
  double best_ms, default_ms;
} fc5_tune;
int fc5_autotune(fc5_ctx *ctx, fc5_band *band, int sample_every, fc5_tune *out);
/* one-call form: create, run, destroy (run_cls5_gpu, fourcore_cls5_gpu_v4.cu:412) */
int fc5_run_units(fc5_ctx *ctx, const fc5_unit *units, size_t n, int mode, fc5_result *res);
/* counters of one class (2..5) in the last run */
int fc5_last_class_result(fc5_ctx *ctx, int cls, fc5_result *out);
/* hits and solutions of the last run of this context */
int fc5_last_hits(fc5_ctx *ctx, fc5_hit *out, size_t cap, size_t *n);
int fc5_last_solutions(fc5_ctx *ctx, fc5_solution *out, size_t cap, size_t *n);
/* fc5_run_units (mode FC5_MODE_BATCHED) for callers that cannot allocate.
 * No counterpart in the reference: it is for the callers that cannot call
 * hipMalloc -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid device allocation.  Every device buffer beyond the inputs and outputs
 * (packed units, launch segment, d table, counters, work-queue and hit counters,
 * hit buffer) lives in one static device array of FC5_NOALLOC_BYTES per GPU,
 * held until the process ends; the call allocates no device memory and no
 * pinned host memory (its host staging is set up once, when a context is
 * created on the device; unit packing uses ordinary host temporaries).
 * Inputs: class-5 Stage-1 units (a line the packing refuses is skipped and
 * counted; any other irregular line is refused with FC5_E_ARG), at most
 * FC5_NOALLOC_UNITS of them, B <= FC5_NOALLOC_BMAX.  Outputs, in caller
 * buffers: the counters (res), the maybe-hits sorted as fc5_last_hits (up to
 * hit_cap; *n_hits gets the number found, at most FC5_NOALLOC_HITS: beyond that
 * the rest are dropped, counted in hits_dropped, and FC5_E_OVERFLOW is
 * returned), and with the option "verify" the solutions (up to sol_cap;
 * *n_sols gets the count) from an exact check of the hits on the host.  Calls
 * on one GPU share its array and are serialized by a lock in the process;
 * the counters are cleared at the start of every call. */
#define FC5_NOALLOC_UNITS 65536
#define FC5_NOALLOC_HITS 65536
#define FC5_NOALLOC_BMAX 5505024 /* free term d = (B - 1) / 21 below 2^18 */
#define FC5_NOALLOC_BYTES 7340032 /* 7 MiB */
int fc5_run_units_noalloc(fc5_ctx *ctx, const fc5_unit *units, size_t n, fc5_result *res, fc5_hit *hits,
                          size_t hit_cap, size_t *n_hits, fc5_solution *sols, size_t sol_cap, size_t *n_sols);
/* closed-form calls of a unit list (the counter `calls`, fourcore_cls5_gpu_v4.cu:383-386) */
int fc5_units_calls(const fc5_unit *units, size_t n, uint64_t *calls);

/* exact check of filter survivors on the device (replaces PairRecover + the
 * exact compare of the host verify, fourcore_xor_store.hpp:23-51,
 * fourcore_cls5_gpu_v4.cu:517-522): for each 128-bit target T_i (targets[2i] low,
 * targets[2i+1] high word, T_i < 2^127) every pair a <= b <= hi[i] with a^6 + b^6 = T_i.  out
 * gets (i, a, b) triples (up to cap, sorted), *n_out the total.  Every run's
 * hit verification goes through this search. */
int fc5_pair_search(fc5_ctx *ctx, const uint64_t *targets, const uint32_t *hi, size_t n, uint32_t *out,
                    size_t cap, size_t *n_out);

/* ---- exact sorted pair table (the sorted pair index QuadSumIndex,
 *      quad_sum.hpp:8-45: build = every pair sum sorted, equal_sum_range = the
 *      pairs of one sum), resident on the device ----
 * Entries: every a^6 + b^6 with 1 <= a <= b <= N plus n_extra extra 128-bit
 * values (extra[2i] low, extra[2i+1] high word; planted exact targets), sorted
 * ascending.  Each entry is one 128-bit word value * 2^s + a, s = bit length of
 * N, a = 0 for an extra value; it must fit 128 bits (N <= 262143 when no extra
 * value exceeds 2 N^6).  16 bytes per entry: N = 76428 -> 2.92e9 entries, 46.7 GB
 * (the build holds twice that while it sorts).  With the option "verify_table"
 * = 1 the exact check of every run's maybe-hits looks each target up here
 * instead of running the pair search (same pairs); a hit whose target is an
 * extra value gets FC5_HIT_EXTRA in its flags.  The table's N must be at least
 * the pair bound of every hit (the main table's N). */
#define FC5_HIT_EXTRA 4u
int fc5_exact_build(fc5_ctx *ctx, int N, const uint64_t *extra, size_t n_extra);
int fc5_exact_clear(fc5_ctx *ctx);
/* entries, N and shift s of the exact table (all 0 when there is none) */
int fc5_exact_info(const fc5_ctx *ctx, uint64_t *n_entries, int *N, uint32_t *shift);
/* copy entries [first, first + n) out as (low, high) word pairs */
int fc5_exact_read(fc5_ctx *ctx, uint64_t first, size_t n, uint64_t *out);
/* equal range of each target T_i: (i, a, b) triples for the pairs with b <= hi[i],
 * (i, 0, 0) for an extra value equal to T_i; layout and cap as fc5_pair_search */
int fc5_exact_lookup(fc5_ctx *ctx, const uint64_t *targets, const uint32_t *hi, size_t n, uint32_t *out,
                     size_t cap, size_t *n_out);

/* ---- (6,1,7) Branch-A class-5 twin (solve_617_v1.cu) ----
 * B^6 = u^6 + w^6 + (21d)^6 + 42^6 (c1^6 + c2^6 + c3^6 + c4^6) with u = 14u'
 * (the term not divisible by 3), w = 6v' (the term not divisible by 7) and the
 * grid term 21d (the odd term).  Options: "use_gate" (shared; --no-gate),
 * "twin_tri_gate" (1; 0 = --no-tri-gate), "twin_c4_div" (3 = the reference
 * kernel's c4 lower bound 3 c4^6 >= T, solve_617_v1.cu:587; 4 = the exact
 * bound 4 c4^6 >= T for the largest of four cores), "twin_ref_tiles" (0 =
 * every c4 searched once; 2048 = the reference launch's c4 tiles, which start
 * at lo4 + 2048 z for z < zmax <= 32 and have no end, solve_617_v1.cu:607-608,
 * 961-979, so a c4 at offset k from lo4 is searched min(32, k/2048 + 1) times
 * and its counts and hits repeat as in the reference; other widths for tests),
 * "hit_cap", "verify". */

/* candidate (GpuCand of class 5, solve_617_v1.cu:354-363): base = B^6 - u^6 - w^6
 * (< 2^127), grid d = res[i] + mod1 * k in [1, fmax1], cores <= lim.  mod1 must be
 * 64 and every residue must have the same sixth power mod 64 (true for the
 * generator's candidates); lim < 65536 (B <= 2,353,973, solve_617_v1.cu:146). */
typedef struct {
  uint64_t B, u, w, base_lo, base_hi;
  uint32_t lim, fmax1, mod1, nres1;
  uint32_t res[8];
} fc5_twin_cand;

/* counters (solve_617_v1.cu:386-390): jobs = (cand, d) cells with a non-empty c4
 * window; c4_total / c4_tri_skipped = tri_total / tri_skipped; calls = probe()
 * invocations (pair gate passed); gated = c3 values the pair gate removed;
 * probes = table cell reads (3 per call); maybe = table maybe-hits */
typedef struct {
  uint64_t jobs, c4_total, c4_tri_skipped, calls, gated, probes, maybe;
  uint64_t n_hits, hits_dropped, solutions, cands, items;
  double kernel_ms, wall_ms;
  uint64_t strict_out;  /* option "strict_gate": calls whose key V = T - c4^6 - c3^6 the strict
                           gate rejected (no query made); 0 without the option */
} fc5_twin_result;

/* a maybe-hit (Hit, solve_617_v1.cu:365: c = c3, d = c4, e = grid d) */
typedef struct {
  uint32_t cand, d, c4, c3;
  uint64_t fp;
} fc5_twin_hit;

/* a verified solution "SOLUTION cls=5 B= a1..a7 (anchor=u) gcd=" (solve_617_v1.cu:1096-1108):
 * pair (a, b) from the table, cores c3, c4, grid d; terms sorted ascending */
typedef struct {
  uint64_t B, u, w;
  uint32_t d, c4, c3, a, b, pad;
  int64_t terms[7];
  int64_t gcd;
} fc5_twin_solution;

/* host generator gen_class_cands(cls 5) for one B (solve_617_v1.cu:807-812,885-906):
 * u' ascending, then v' ascending.  Writes up to cap, *n gets the full count. */
int fc5_twin_cands(uint64_t B, double u_lo, double u_hi, fc5_twin_cand *out, size_t cap, size_t *n);
/* run candidates (gpu_run with k_find4_cls5, then expand_xor_hits + verify_hit,
 * solve_617_v1.cu:924-1006,1019-1110) in one persistent launch */
int fc5_twin_run(fc5_ctx *ctx, const fc5_twin_cand *cands, size_t n, fc5_twin_result *res);
/* Branch A class 5 over B in [B_lo, B_hi] (B odd, 3 and 7 not dividing B): the
 * campaign loop of solve_617_v1.cu:1895-1950 for classes = "5" */
int fc5_twin_run_range(fc5_ctx *ctx, uint64_t B_lo, uint64_t B_hi, double u_lo, double u_hi,
                       fc5_twin_result *res);
int fc5_twin_last_hits(fc5_ctx *ctx, fc5_twin_hit *out, size_t cap, size_t *n);
int fc5_twin_last_solutions(fc5_ctx *ctx, fc5_twin_solution *out, size_t cap, size_t *n);

/* ---- resident serving loop (no reference counterpart: the reference runs one
 *      band per process, fourcore_cls5_gpu_v4.cu:596-709).  A server thread owns
 *      the context's stream; bands of Stage-1 units (classes 2-5) are submitted
 *      at any time, admitted as a counter slot and room in the resident unit
 *      arenas free up, and served by continuous batching: every scheduling
 *      round launches the persistent kernel over a slice of every admitted band,
 *      so bands that arrive while others run share the device with them.  A
 *      band's counters, hits and solutions equal those of fc5_band_run on the
 *      same units.  While the server runs, only fc5_serve_* calls (and
 *      fc5_last_error) may be used on the context; others return FC5_E_STATE.
 *      The pair table (and side table; with "verify_table" the exact table) must
 *      be loaded before fc5_serve_start;
 *      bands whose B exceed what the table's N covers are refused. ---- */
typedef struct {
  int max_bands;          /* bands admitted at once (counter slots), 1..64; default 16 */
  int rows_per_item;      /* 4, 8, 16, 32; default: the context option, or 16 when that is 0 */
  uint64_t arena_units5;  /* resident class-5 unit capacity (64 B each); default 1<<23 */
  uint64_t arena_units3;  /* resident class 2-4 unit capacity (128 B each); default 1<<22 */
  double round_points;    /* grid points per scheduling round; default 2e10 (~30 ms) */
  uint64_t mem_budget;    /* bytes: table + arenas + run buffers must fit; default 288e9 */
} fc5_serve_opts;
typedef struct {
  double t_submit, t_admit, t_done;  /* seconds on the server's steady clock */
  uint64_t rounds;                   /* scheduling rounds the band took part in */
  double t_release;                  /* its last round ended: counter slot and arena ranges
                                        given back (the exact check of its hits follows) */
} fc5_serve_timing;
typedef struct {
  uint64_t submitted, admitted, done, pending, active, rounds;
  uint64_t arena5_used, arena3_used, device_bytes;
  double busy_ms;  /* summed event time of the rounds */
} fc5_serve_stats;
int fc5_serve_start(fc5_ctx *ctx, const fc5_serve_opts *opts); /* opts may be NULL */
/* copy in and pack a band; returns its id.  cls may be NULL (all class 5).
 * Lines of classes 2-5 must be Stage-1 units (a line that only the per-point
 * paths of fc5_band_run can take is refused with FC5_E_ARG); lines of other
 * classes and lines the packing refuses are skipped and counted */
int fc5_serve_submit(fc5_ctx *ctx, const uint32_t *cls, const fc5_unit *units, size_t n, uint64_t *band_id);
/* FC5_OK and the band's result when done, FC5_E_PENDING otherwise; a finished
 * band some of whose hits did not fit a round's hit buffer (hit_cap records,
 * shared by the bands of the round) returns FC5_E_OVERFLOW with the result:
 * n_hits counts the hits kept, hits_dropped the band's own hits that were
 * dropped (the kernel counts every band's hit records, so bands of the same
 * round whose hits were all kept are not flagged) */
int fc5_serve_poll(fc5_ctx *ctx, uint64_t band_id, fc5_result *res);
/* blocks until done or timeout_s elapses (timeout_s < 0: no limit) */
int fc5_serve_wait(fc5_ctx *ctx, uint64_t band_id, double timeout_s, fc5_result *res);
int fc5_serve_class_result(fc5_ctx *ctx, uint64_t band_id, int cls, fc5_result *out);
int fc5_serve_hits(fc5_ctx *ctx, uint64_t band_id, fc5_hit *out, size_t cap, size_t *n);
int fc5_serve_solutions(fc5_ctx *ctx, uint64_t band_id, fc5_solution *out, size_t cap, size_t *n);
int fc5_serve_timing_of(fc5_ctx *ctx, uint64_t band_id, fc5_serve_timing *out);
/* submit the Stage-1 units of classes cls_mask (bits 2..5) for B in [lo, hi],
 * emitted on the device (on a stream of their own, beside the server's rounds) */
int fc5_serve_submit_range(fc5_ctx *ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac,
                           double hi_frac, uint64_t *band_id);
/* forget a finished band (its results are freed) */
int fc5_serve_release(fc5_ctx *ctx, uint64_t band_id);
int fc5_serve_stats_get(fc5_ctx *ctx, fc5_serve_stats *out);
/* finish every submitted band, then stop the server thread */
int fc5_serve_stop(fc5_ctx *ctx);

/* ---- measurement (no reference counterpart: the roofline microbenchmarks) ---- */
/* n pure xor-filter queries of pseudo-random keys against the main table at
 * full occupancy (probe_ms), and 3n raw random 8-byte loads of its cells
 * (raw_ms), both under the context's "load_policy" */
int fc5_measure_gathers(fc5_ctx *ctx, uint64_t n, double *probe_ms, double *raw_ms);
/* a streaming device copy of `bytes` bytes (reads and writes that many) */
int fc5_measure_copy(fc5_ctx *ctx, uint64_t bytes, double *ms);
/* n raw random 8-byte loads confined to the first region_cells cells of the
 * main table (the L2 / memory-side-cache / HBM curve of random gathers) */
int fc5_measure_gathers_region(fc5_ctx *ctx, uint64_t region_cells, uint64_t n, double *ms);
/* a read-only streaming pass over a fresh device buffer of `bytes` bytes */
int fc5_measure_read(fc5_ctx *ctx, uint64_t bytes, double *ms);
/* integer VALU issue rate in wavefront instructions per second at full
 * occupancy: kind 0 = 32-bit (v_xad_u32), 1 = v_mad_u64_u32, 2 = 128-bit
 * add-with-carry chains, 3 = one v_mad_u64_u32 per 13 32-bit operations */
int fc5_measure_valu(fc5_ctx *ctx, int kind, double *instr_per_s);
/* partitioned probing of n <= 2^32-1 pseudo-random keys (those of
 * fc5_measure_gathers) against the main table: keys radix-partitioned by table
 * slice, each slice's low 8 cell bits in LDS, candidates fully queried.
 * ms[9]: first partition, fine partition and lookup of cells 0, 1, 2 (six
 * stages), full queries, then the direct queries of the same keys.
 * counts[8]: accepted and the sum of their hashes, the same for the direct
 * queries, full queries made, records lost (must be 0), slices per third,
 * coarse buckets per third */
int fc5_measure_partitioned(fc5_ctx *ctx, uint64_t n, double *ms, uint64_t *counts);
/* the last batched run with partitioned probing (option probe_mode):
 * ms[11]: host-timed emit launches, partition pipelines and identify launch,
 * then the pipeline's stages timed by events and summed over the pieces
 * (first partition, fine partition and lookup of cell 0, of cell 1, fine
 * partition and last lookup of cell 2, full queries);
 * counts[4]: pieces, accepted hashes, full queries, emit chunks filled */
int fc5_part_stats(fc5_ctx *ctx, double *ms, uint64_t *counts);

/* ---- host helpers shared by the CLIs (no device use) ---- */
/* Stage-1 emission over [lo,hi] for classes in cls_mask (bit c = class c),
 * u-band [lo_frac,hi_frac) (fourcore_hunt_v3.cpp:371-381, fourcore_classes_v3.hpp:137-154).
 * Writes up to cap triples; *n gets the full count. */
int fc5_emit_units(uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac, double hi_frac,
                   uint32_t *out_cls, uint64_t *out_B, uint64_t *out_u, size_t cap, size_t *n);
/* ---- Stage-1 on the device (no reference counterpart: the reference emits
 *      units on the host, fourcore_hunt_v3.cpp:371-381) ---- */
/* the same units, in the same order, as fc5_emit_units for classes 2-5
 * (cls_mask bits 2..5; class 1 stays on the host) */
int fc5_emit_units_device(fc5_ctx *ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac, double hi_frac,
                          uint32_t *out_cls, uint64_t *out_B, uint64_t *out_u, size_t cap, size_t *n);
/* a band straight from a B range: Stage-1 on the device, then fc5_band_create_mixed */
int fc5_band_create_range(fc5_ctx *ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac, double hi_frac,
                          fc5_band **out);
/* ---- host helpers, continued ---- */
/* host-only xor pair-table build + save (xor_build_save.cpp:21-57) */
int fc5_build_pairs_file(int N, int r, const char *path, fc5_store_header *hdr_out);

#ifdef __cplusplus
}
#endif
#endif /* FC5_H */
