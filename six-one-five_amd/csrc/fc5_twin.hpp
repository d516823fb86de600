// fc5_twin.hpp -- device layouts of the (6,1,7) class-5 twin search (gfx950).
//
// Candidate (B, u = 14u', w = 6v') with base = B^6 - u^6 - w^6 and the grid
// term 21d, d = res[i] + 64k <= fmax1.  Every d of the residue classes has
// the same d^6 mod 64, so R = base - (21d)^6 is a multiple of 42^6 and
//     T(d) = R / 42^6 = Q - floor(d^6 / 64),   Q = floor(base / 42^6),
// which the device evaluates in 128 bits.  The search is T = c1^6 + ... + c4^6
// with c4 in [lo4, min(iroot6(T), lim)] (lo4 = least c with div * c^6 >= T),
// then c3 in [least c with 3 c^6 >= T - c4^6, min(iroot6(T - c4^6), lim, c4)],
// then the pair (c1, c2) through the xor table with key (T - c4^6 - c3^6) mod 2^64.
#pragma once

#include "fc5_kernels.hpp"

namespace fc5 {

struct DevTwinCand {
  u64 Q_lo, Q_hi;       // floor(base / 42^6)
  u32 fmax1, lim;
  u32 rows, item_base;  // rows = nres * (fmax1/64 + 1); first work item
  u32 uid, nres;        // index in the caller's candidate list
  u8 res[8];            // d residues mod 64
  u32 pad[4];
};
static_assert(sizeof(DevTwinCand) == 64, "DevTwinCand is 64 bytes");

struct TwinCounters {
  u64 jobs, c4_total, c4_tri_skipped, calls, gated, maybe;
  u64 queries;  // strict gate on: calls that pass it and are queried
};

struct TwinHit {
  u32 cand, d, c4, c3;
  u64 fp;
};
static_assert(sizeof(TwinHit) == 24, "TwinHit is 24 bytes");

struct TwinArgs {
  const DevTwinCand* cands;
  u32 n_cands, n_items;
  const u32* gate;  // 124488-bit pair-sum bitmap (g504 && g247)
  const u32* tri;   // triple-sum sets: 504-bit map in words 0..15, 247-bit map in 16..23
  u32 use_gate, use_tri;
  u32 c4_div;       // 3: the reference kernel's c4 bound; 4: exact bound
  u32 tile;         // 0: every c4 once; > 0: the reference launch's unbounded c4 tiles of this width
  u32 strict, pad_; // strict: the strict residue gate (fc5_strict.hpp) on V = T - c4^6 - c3^6 before each query
  const u32* scsieve;  // strict: the strict sieve table over consecutive c (strict_csieve_table)
  XorDev main;
  TwinHit* hits;
  u64 hit_cap;
  u64* nhits;
  TwinCounters* ctr;
  u32* item_ctr;
};

// reference-shaped candidate (GpuCand class 5, solve_617_v1.cu:354-363)
struct PlainTwinCand {
  u64 b_lo, b_hi;  // base = B^6 - u^6 - w^6
  u32 lim, fmax1, nres, uid;
  u32 res[8];
};
static_assert(sizeof(PlainTwinCand) == 64, "PlainTwinCand is 64 bytes");

struct PlainTwinArgs {
  const PlainTwinCand* cands;
  u32 n_cands, use_gate, use_tri, c4_div;
  u32 tile, pad;    // as TwinArgs::tile (0: bounded tiles of 2048, each c4 once)
  GateWords gate;   // pair-sum sets mod 504 and 247
  const u32* tri;   // as TwinArgs::tri
  const u16* p6;    // x^6 mod 504 (504 entries) then mod 247 (247 entries)
  u64 inv216, c_m504, c_m247, inv42_247;
  XorDev main;
  TwinHit* hits;
  u64 hit_cap;
  u64* nhits;
  TwinCounters* ctr;
};

int launch_twin_batched(const TwinArgs& a, int grid, hipStream_t s);  // a.strict picks the variant
int twin_blocks_per_cu(bool strict);
// the reference launch shape: grid (candidate, d-row, c4 tile of 2048 or a.tile)
int launch_twin_plain(const PlainTwinArgs& a, u32 ymax, u32 zmax, hipStream_t s);

}  // namespace fc5
