// fc5_part.hpp -- partitioned xor-filter probing: layouts and launchers
// (kernels in fc5_part.hip).
#pragma once

#include "fc5_kernels.hpp"

namespace fc5 {

// The lookups use the low kLowBits bits of each cell: a slice of 65,536 cells
// of one third fits LDS in 64 KiB, and a record's partial xor is one byte.
using low_t = u8;
constexpr u32 kLowBits = 8;
constexpr u32 kSliceShift = 16;                 // slice = 65536 cells of one third
constexpr u32 kSliceCells = 1u << kSliceShift;  // staged in LDS as low_t: 64 KiB
constexpr u32 kFineShift = 8;                   // 256 fine buckets (slices) per coarse bucket
constexpr u32 kPartMaxBins = 512;               // coarse buckets per third at most
constexpr u32 kCoarseFillStride = 64;           // coarse fill counters 256 bytes apart
constexpr u32 kCoarseSubs = 32;                 // regions per coarse bucket

struct PartGeom {
  XorDev x;          // the table: block, divisor, expanded cells (full queries)
  const low_t* lo;   // low kLowBits bits of every cell, each third padded to whole slices
  u64 lo_third;      // cells per third in lo (ns slices: the slice (p, f) starts at (p ns + f) S)
  u32 slice_shift;   // log2 of the slice length
  u32 fine_shift;    // log2 of the fine buckets per coarse bucket
  u32 nc;            // coarse buckets per third
  u32 mask_lo;       // low kLowBits bits of the fingerprint mask
};

// nb buckets, each split into `subs` regions of cap records (region
// bucket * subs + sub): hashes, partial xors (nullptr: zero, first cell) and
// one fill counter per region (fill_stride u32 apart; a fill past cap means
// the excess went to the side list).  Producers spread over the regions of a
// bucket so that no counter takes more than a share of the reservations: one
// counter serialises its atomics (about 35 ns each, measured).
struct PartSet {
  u64* h;
  low_t* a;
  u32* fill;
  u64 cap;
  u32 nb, subs, fill_stride;
};

// hashes that get the full query: candidates and records past a bucket's
// capacity; lost counts pushes past cap (the caller must treat it as an error)
struct PartSide {
  u64* h;
  u32* n;
  u32* lost;
  u32 cap;
};

// geometry for a table of `block` cells per third
inline PartGeom make_part_geom(const XorDev& x, const low_t* lo) {
  PartGeom g{};
  g.x = x;
  g.lo = lo;
  g.slice_shift = kSliceShift;
  g.fine_shift = kFineShift;
  const u64 ns = (x.block + kSliceCells - 1) >> kSliceShift;
  g.lo_third = ns << kSliceShift;
  g.nc = (u32)((ns + (1u << kFineShift) - 1) >> kFineShift);
  g.mask_lo = (u32)(x.mask & ((1u << kLowBits) - 1u));
  return g;
}

// the low-bits table of g (3 * g.lo_third entries) from the expanded cells of g.x
int launch_part_low(const PartGeom& g, low_t* lo, hipStream_t s);
// n hashes at h (tile_fill: tile t holds tile_fill[t] <= 4096 hashes at h + 4096 t)
int launch_part_first(const u64* h, u64 n, const u32* tile_fill, const PartGeom& g, const PartSet& out,
                      const PartSide& side, hipStream_t s);
int launch_part_fine(int pass, const PartGeom& g, const PartSet& in, const PartSet& out, const PartSide& side,
                     hipStream_t s);
int launch_part_lookup(int pass, const PartGeom& g, const PartSet& in, const PartSet& out, const PartSide& side,
                       hipStream_t s);
int launch_part_last(const PartGeom& g, const PartSet& in, const PartSide& side, hipStream_t s);
// full queries of side.h[0, min(*side.n, side.cap)): accepted hashes appended to acc
int launch_part_check(const PartGeom& g, const PartSide& side, const PartSide& acc, int grid, hipStream_t s);
// hashes of the pseudo-random keys of k_probe_only (salt), and their direct
// queries: count and sum of the accepted hashes
int launch_part_gen(const XorDev& x, u64 n, u64 salt, u64* h, hipStream_t s);
int launch_part_ref(const XorDev& x, const u64* h, u64 n, unsigned long long* cnt_sum, int grid, hipStream_t s);
// count and sum of h[0, min(*n, cap))
int launch_part_sum(const u64* h, const u32* n, u32 cap, unsigned long long* cnt_sum, hipStream_t s);

}  // namespace fc5
