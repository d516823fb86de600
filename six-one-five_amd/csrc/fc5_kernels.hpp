// fc5_kernels.hpp -- device data layouts and kernel launchers (gfx950).
//
// Layouts in HBM:
//   pair table   one u64 per xor cell ("expanded"; the r-bit packed file
//                layout is widened once at upload), m_cells * 8 bytes
//   gate         124488-bit bitmap of T mod 504*247 (g504 && g247), 15.2 KiB,
//                staged into LDS by every workgroup
//   ndg table    [64][K] u32: (124488 - floor(d^6/64) mod 124488) for
//                d = r + 64k, transposed by d mod 64 so that the lanes of a
//                wave read consecutive k of one residue class
//   units        DevUnit5, 64 bytes each, T-space form (see fc5_host.hpp)
//   hits         HitRec, 24 bytes, append-only with a capacity
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fc5_moddiv.hpp"

namespace fc5 {

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using u128 = unsigned __int128;


struct XorDev {
  const u64* cells;  // expanded cells
  u64 block;         // m_cells / 3
  u64 mask;          // low r bits
  u64 seed;          // mix seed
  ModDiv md;
  u32 present;       // 0: queries answer "no" (empty table), as block == 0
  u32 policy;        // cell load cache policy: 0 default, 1 nontemporal,
                     // 2 device-scope coherent (sc1), 3 system-scope (sc0 sc1)
};

// one cell load under the table's cache policy (uniform branch)
__device__ __forceinline__ u64 load_cell_policy(const u64* p, u32 policy) {
  switch (policy) {
    case 1: return __builtin_nontemporal_load(p);
    case 2: return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    case 3: return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    default: return *p;
  }
}

struct DevUnit5 {
  u64 Q_lo, Q_hi;       // T = Q - floor(e^6/729) - floor(d^6/64)
  u32 fmax1, fmax2;
  u32 rows, item_base;  // rows = 6 * (fmax1/729 + 1); first work item
  u32 te, td;           // e^6 mod 729, d^6 mod 64
  u32 uid, band;        // index in the caller's unit list; counter slot
  u16 res1[6];          // e residues mod 729
  u8 res2[4];           // d residues mod 64
};
static_assert(sizeof(DevUnit5) == 64, "DevUnit5 is 64 bytes");

// class 2-4 unit in T-space: T(d) = Q - floor(d^6/mod) for d = res[i] + mod*k,
// mod = 729 (cls2, f=14), 64 (cls3, f=21), 46656 (cls4, f=7); f^6 * mod = 42^6
struct DevUnit3 {
  u64 Q_lo, Q_hi;
  u64 inv_lo, inv_hi;   // inverse of the odd part of mod, mod 2^128 (exact quotients)
  u32 fmax, lim;
  u32 rows, item_base;  // rows = nres * (fmax/mod + 1); first work item (numbered per arena)
  u32 mod, dr;          // dr = d^6 mod mod over the residue classes
  u32 shift, nres;      // mod = 2^shift * odd
  u32 uid, band;
  u16 res[24];
  u32 pad[2];
};
static_assert(sizeof(DevUnit3) == 128, "DevUnit3 is 128 bytes");

// reference-shaped unit (Unit5D, fourcore_cls5_gpu_v4.cu:346-351)
struct PlainUnit {
  u64 b0, b1, b2;
  u32 fmax1, fmax2, mod1, mod2, nres1, nres2, uid, pad;
  u32 res[24];
};

struct Counters {
  u64 calls, neg, zero, gated, probes, maybe, side_maybe, pos, jobs, failT;
  u64 recs;     // hit records produced (written or dropped at hit_cap), batched kernel
  u64 queries;  // strict gate on: xor queries made (gate survivors that pass the strict gate)
};

struct HitRec {
  u32 unit, d, e, flags;  // flags: bit0 main table, bit1 side table; batched kernel: counter slot << 8
  u64 fp;
};

struct GateWords {  // reference bitmaps (GateData.g504/g247, fourcore_cls5_gpu_v4.cu:111-116)
  u64 g504[8];
  u64 g247[4];
};

// A run of consecutive work items of one launch: items g0.. map to the
// unit-local items item0.. of the units [u0, u0 + n_units) of one arena
// (kind 0: class-5 units, kind 1: class 2-4 units).
struct Segment {
  u32 g0, item0, u0, n_units, kind, pad;
};

struct BatchedArgs {
  const DevUnit5* units;  // class-5 unit arena (ring entries index it)
  const Segment* segs;
  u32 n_segs, total;      // segments and work items of this launch
  const u32* gate;  // 124488-bit bitmap
  const u32* ndg;   // [64][ndg_K]
  u32 ndg_K, use_gate;
  u32 strict;  // the strict residue gate before every xor query
  u32 sieve;   // class-5 sweep: 1 bit sieve over d (fc5_sieve.hpp), 0 one LDS bitmap test per point
  XorDev main, side;
  HitRec* hits;
  u64 hit_cap;
  u64* nhits;     // hit slot counter (shared by all kernels of a run)
  Counters* ctr;  // indexed by DevUnit5::band
  u32* item_ctr;
  u64 inv729;  // 729^-1 mod 2^64
  const DevUnit3* units3;  // class 2-4 unit arena
  u32 grab, pad_;           // work items per work-counter update (set by launch_cls5_batched)
  // Probe mode (fc5_part.hip).  0: every probe is an xor query of the main
  // table.  1 (emit): each probe's hash h = mix64(key + seed) is appended to
  // the wavefront's chunk of `flat` (chunk c holds chunk_fill[c] hashes at
  // c * kEmitChunk); no hit records; a probe that finds no free chunk is
  // queried at once and, if accepted, its hash pushed to acc_h; a wavefront
  // takes no work item once its latest chunk is >= chunk_limit.  2 (identify): the main table's
  // answer for a probe is whether h is in the sorted set set_h[0, set_n)
  // (set_bits: 2^kSetFilterBits-bit filter of the hashes' top bits).  Modes 1 and 2
  // need the strict gate off.
  u32 pmode, chunk_limit;
  u64* flat;
  u32* chunk_ctr;
  u32* chunk_fill;
  u32 n_chunks, set_n;
  u64* acc_h;
  u32* acc_n;
  u32* acc_lost;
  u32 acc_cap, pad2_;
  const u64* set_h;
  const u32* set_bits;
};
constexpr u32 kEmitChunk = 4096;  // hashes per emit chunk (one partition tile)
// probe mode 2: the accepted set's filter, 2^20 bits (128 KiB, L2-resident): a
// probe passes it with odds (accepted hashes) / 2^20 before the binary search
constexpr u32 kSetFilterBits = 20;

// reference-shaped find3 candidate (Cand, fourcore_find_v4.cu:102-109)
struct PlainCand3 {
  u64 q_lo;
  u32 lo, hi;
  u32 r504, r247;
  u32 uid, d;
  u32 band, srow;  // srow, srow1: strict-gate row words of the candidate's T (strict_row)
  u32 srow1, pad;
};

struct PlainF3Args {
  const PlainCand3* cands;
  int n;
  int use_gate;
  int strict, pad_;
  GateWords gate;
  const u16* p6;  // [504] x^6 mod 504, then [247] x^6 mod 247 (GateData.p6504/p6247)
  XorDev main, side;
  HitRec* hits;
  u64 hit_cap;
  u64* nhits;
  Counters* ctr;  // indexed by band
};

struct PlainArgs {
  const PlainUnit* units;
  int n_units;
  int use_gate;
  int strict, pad_;
  GateWords gate;
  XorDev main, side;
  u64 inv216, c_m504, c_m247, inv42_247;
  HitRec* hits;
  u64 hit_cap;
  u64* nhits;
  Counters* ctr;
};

// exact pair search (fc5_verify.hip): pairs a <= b <= hi with a^6 + b^6 = T
struct PairJob {
  u64 T_lo, T_hi;
  u32 hi, pad;
};
struct PairHit {
  u32 job, a, b, pad;
};

// Stage-1 emission on the device (fc5_stage1.hip), classes 2-5
struct Stage1Args {
  u64 lo, count;  // B = lo .. lo + count - 1
  double lo_frac, hi_frac;
  u32 cls_mask, pad;
  const u64* z2;  // the 4 odd sixth roots of unity mod 2^6
  const u64* z3;  // the 6 mod 3^6
  const u64* z7;  // the 6 mod 7^6
  u64 e2_14, e7_14, e3_21, e7_21;  // CRT coefficients for 14^6 = 2^6 7^6 and 21^6 = 3^6 7^6
  u64* counts;           // pass 0: units per B
  const u64* offsets;    // pass 1: exclusive prefix of counts
  u32* out_cls;
  u64 *out_B, *out_u;
};

constexpr int kRowsPerItem = 16;  // rows (e- or d-rows) per work item of a serving round, unless fixed

// launchers (return hipError_t as int)
// rows: 4, 8, 16, 32; a.strict picks the variant with the strict gate stage
int launch_cls5_batched(const BatchedArgs& a, int grid, int rows, hipStream_t s);
// the strict sieve tables of the batched kernel into the current device (once per device)
int upload_sieve_tables();
// the current device's strict sieve table over consecutive c (for the twin kernel)
const u32* strict_csieve_table();
int launch_cls5_plain(const PlainArgs& a, int n_units, int ymax, hipStream_t s);
int launch_find3_plain(const PlainF3Args& a, hipStream_t s);
int launch_expand_cells(const u32* packed_words, u64 m_cells, u32 r, u64* cells, hipStream_t s);
int launch_query(const XorDev& x, const u64* keys, u64 n, u8* out, hipStream_t s);
int launch_fill_u64(u64* p, u64 n, u64 v, hipStream_t s);
int batched_blocks_per_cu(int rows, bool strict, bool sieve);  // occupancy limit of the kernel variant
int launch_probe_only(const XorDev& x, u64 n, u64 salt, unsigned long long* sink, int grid, hipStream_t s);
int launch_gather_raw(const u64* cells, u64 m, u64 n, u64 salt, u32 policy, unsigned long long* sink, int grid,
                      hipStream_t s);
int launch_copy_u64(const u64* src, u64* dst, u64 n, int grid, hipStream_t s);
int launch_read_stream(const void* src, u64 bytes, unsigned long long* sink, int grid, hipStream_t s);
// integer VALU issue microbenchmark; kind 0..3, VALU instructions per loop step per wave: kValuInstrs[kind]
constexpr u32 kValuInstrs[4] = {8, 8, 32, 112};
int launch_valu(int kind, u32 iters, u32 seed, unsigned long long* sink, int grid, hipStream_t s);
int launch_stage1(const Stage1Args& a, int pass, hipStream_t s);
// exclusive prefix sum of n u64 (temp == nullptr: query temp_bytes)
int stage1_scan(const u64* counts_in, u64* offsets_out, u64 n, void* temp, size_t& temp_bytes, hipStream_t s);
int launch_pair_search(const PairJob* jobs, u32 n, PairHit* out, u32 cap, u32* count, hipStream_t s);
// exact sorted pair table (fc5_exact.hip): rows of entries (a^6 + b^6) * 2^s + a
// for 1 <= a <= b <= N; radix sort of n entries on bits [0, end_bit) (temp ==
// nullptr: query temp_bytes); per-target equal ranges (PairHit.pad = 1: extra value)
int launch_exact_rows(u128* keys, u32 N, u32 s, hipStream_t st);
int exact_sort(u128* keys, u128* alt, u64 n, u32 end_bit, void* temp, size_t& temp_bytes, hipStream_t st,
               bool& sorted_in_alt);
int launch_exact_lookup(const u128* keys, u64 n, u32 s, const PairJob* jobs, u32 nj, PairHit* out, u32 cap,
                        u32* count, hipStream_t st);

}  // namespace fc5
