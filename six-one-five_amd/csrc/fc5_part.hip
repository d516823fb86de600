// fc5_part.hip -- partitioned xor-filter probing (gfx950).
//
// A filter query reads three cells at hashed positions of a table far larger
// than every cache (one cell in each third of it, spike/ribbon/xor_filter.hpp:
// 53-64), and each of those 8-byte gathers fills a whole 128-byte L2 line.
// Here a batch of queries is taken to the table instead: the queries' hashes
// are radix-partitioned, in two levels, by the slice (kSliceCells cells of one
// third) that holds the current cell; a workgroup stages its slice's low 8
// cell bits in LDS, looks each query of its bucket up there, and the query's
// 8-bit partial xor travels with its hash into the partition of the next
// cell.  After the third cell a query whose partial xor equals the low 8
// bits of its fingerprint is a candidate (about 2^-8 of the rest); every
// candidate gets the full query of all r bits against the expanded cells, so
// the accepted set is exactly the filter's.  A record that finds its bucket
// full also gets the full query at once (nothing is dropped).
//
// Records: the 64-bit hash h = mix64(key + seed) and the partial xor (a byte),
// in separate arrays; buckets have a fixed capacity, so a partition is one
// pass (per-tile histogram in LDS, one reservation atomic per non-empty
// bucket and tile, runs written contiguously).
//
// k_part_low     the low 8 bits of every cell (the lookups' table)
// k_part_first   hashes (in any order) -> coarse buckets of the first cell
// k_part_fine    coarse bucket -> fine buckets (= slices) of the same cell
// k_part_lookup  fine bucket: slice in LDS, xor, coarse buckets of the next cell
// k_part_last    fine bucket of the third cell: slice in LDS, candidates
// k_part_check   full queries of candidates and overflowed records
// k_part_gen, k_part_ref   the hashes of k_probe_only's pseudo-random keys
//                and their direct queries (the measurement of this scheme)
#include <hip/hip_runtime.h>

#include "fc5_device.hpp"
#include "fc5_part.hpp"

namespace fc5 {

// 4096 records per tile, so a coarse (<= 512 buckets) or fine (256 buckets)
// partition writes runs of about 8 to 30 records per bucket and tile
constexpr int kPartThreads = 256;
constexpr int kLookThreads = 1024;
constexpr int kTileRecs = 4096;

// LDS of one tile scatter into at most NB buckets (NB = 256: 47 KiB, three
// blocks per CU)
template <int NB>
struct TileSmem {
  u64 h[kTileRecs];
  low_t a[kTileRecs];
  u16 d[kTileRecs];
  u32 hist[NB];
  u32 lofs[NB];
  u32 gbase[NB];
  u32 total;
};

// cell index of pass p (0, 1, 2) within its third of the table
__device__ __forceinline__ u64 seg_cell(u64 h, int p, const XorDev& x) {
  const u64 v = p == 0 ? h : mix64(h ^ (p == 1 ? kCellSalt1 : kCellSalt2));
  return mod_fast(v, x.md);
}

// full query of a hash (d_query after the key's mix)
__device__ __forceinline__ bool query_hash(const XorDev& x, u64 h) {
  u64 c[3];
  xor_cells(h, x.block, x.md, c);
  return (x.cells[c[0]] ^ x.cells[c[1]] ^ x.cells[c[2]]) == (h & x.mask);
}

__device__ __forceinline__ void side_push(const PartSide& s, u64 h) {
  const u32 k = atomicAdd(s.n, 1u);
  if (k < s.cap)
    s.h[k] = h;
  else
    atomicAdd(s.lost, 1u);
}

// Thread t's records j < R of a tile of cnt >= 1 records (positions t + NT j;
// positions past the end read the last record, so every load is issued
// before the first use and none is out of bounds); src == nullptr: zeros
template <int NT, int R, class T>
__device__ __forceinline__ void load_tile(T (&v)[R], const T* __restrict__ src, u32 cnt) {
  if (!src) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = 0;
    return;
  }
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = src[min(threadIdx.x + NT * j, cnt - 1)];
}

// Scatter one tile's records into region `sub` of out's buckets bucket0 +
// dig[j] (dig[j] < nbins <= kPartMaxBins).  Thread t holds records j < R, bit
// j of `valid` set for the real ones.  Through LDS, so each bucket's records
// leave as one contiguous run.  Every thread of the block must call it.
template <int NT, int R, int NB>
__device__ __forceinline__ void tile_scatter(TileSmem<NB>& s, const u64 (&h)[R], const low_t (&a)[R],
                                             const u32 (&dig)[R], u32 valid, u32 nbins, u32 bucket0, u32 sub,
                                             const PartSet& out, const PartSide& side) {
  static_assert(NT * R == kTileRecs, "tile shape");
  const u32 t = threadIdx.x;
  for (u32 b = t; b < nbins; b += NT) s.hist[b] = 0;
  __syncthreads();
  u32 rank[R];
#pragma unroll
  for (int j = 0; j < R; ++j) rank[j] = ((valid >> j) & 1u) ? atomicAdd(&s.hist[dig[j]], 1u) : 0u;
  __syncthreads();
  if (t < 64) {  // exclusive scan of the histogram, NB / 64 bins per lane at most
    const u32 per = (nbins + 63) / 64;
    u32 sum = 0;
    for (u32 k = 0; k < per; ++k) {
      const u32 b = t * per + k;
      if (b < nbins) sum += s.hist[b];
    }
    u32 incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
      const u32 v = (u32)__shfl_up((int)incl, o, 64);
      if ((int)t >= o) incl += v;
    }
    u32 run = incl - sum;
    for (u32 k = 0; k < per; ++k) {
      const u32 b = t * per + k;
      if (b < nbins) {
        s.lofs[b] = run;
        run += s.hist[b];
      }
    }
    if (t == 63) s.total = incl;
  }
  for (u32 b = t; b < nbins; b += NT) {
    const u32 c = s.hist[b];
    s.gbase[b] = c ? atomicAdd(out.fill + ((u64)(bucket0 + b) * out.subs + sub) * out.fill_stride, c) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < R; ++j)
    if ((valid >> j) & 1u) {
      const u32 pos = s.lofs[dig[j]] + rank[j];
      s.h[pos] = h[j];
      s.a[pos] = a[j];
      s.d[pos] = (u16)dig[j];
    }
  __syncthreads();
  const u32 total = s.total;
  for (u32 i = t; i < total; i += NT) {
    const u32 b = s.d[i];
    const u64 k = (u64)s.gbase[b] + (i - s.lofs[b]);
    const u64 at = ((u64)(bucket0 + b) * out.subs + sub) * out.cap + k;
    if (k < out.cap) {
      out.h[at] = s.h[i];
      if (out.a) out.a[at] = s.a[i];
    } else {
      side_push(side, s.h[i]);
    }
  }
  __syncthreads();  // s is reused by the block's next tile
}

// cell c of third p -> lo[p lo_third + c]; the padding past each third stays unset
__global__ void k_part_low(const u64* __restrict__ cells, u64 block, u64 lo_third, low_t* __restrict__ lo) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * block; i += (u64)gridDim.x * blockDim.x) {
    const u64 p = i / block;
    lo[p * lo_third + (i - p * block)] = (low_t)cells[i];
  }
}

// one block per tile of the input hashes: n hashes from src, or, with
// tile_fill, tile t holds tile_fill[t] hashes at src + t * kTileRecs
template <int NB>
__global__ __launch_bounds__(kPartThreads) void k_part_first(const u64* __restrict__ src, u64 n,
                                                              const u32* __restrict__ tile_fill, PartGeom g,
                                                              PartSet out, PartSide side) {
  constexpr int NT = kPartThreads, R = kTileRecs / NT;
  __shared__ TileSmem<NB> s;
  const u64 base = (u64)blockIdx.x * kTileRecs;
  const u32 cnt = tile_fill ? min(tile_fill[blockIdx.x], (u32)kTileRecs)
                            : (base < n ? (u32)min((u64)kTileRecs, n - base) : 0u);
  if (cnt == 0) return;
  u64 h[R];
  low_t a[R];
  u32 dig[R];
  u32 valid = 0;
  load_tile<NT, R>(h, src + base, cnt);
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const bool ok = threadIdx.x + NT * j < cnt;
    a[j] = 0;
    dig[j] = ok ? (u32)(seg_cell(h[j], 0, g.x) >> (g.slice_shift + g.fine_shift)) : 0u;
    valid |= (ok ? 1u : 0u) << j;
  }
  tile_scatter<NT, R, NB>(s, h, a, dig, valid, g.nc, 0, blockIdx.x % out.subs, out, side);
}

// grid (coarse regions, tiles): coarse bucket -> its 2^fine_shift fine
// buckets; the regions vary fastest, so the blocks in flight reserve in the
// fine buckets of many coarse buckets at once
__global__ __launch_bounds__(kPartThreads) void k_part_fine(int p, PartGeom g, PartSet in, PartSet out,
                                                             PartSide side) {
  constexpr int NT = kPartThreads, R = kTileRecs / NT;
  __shared__ TileSmem<1 << kFineShift> s;
  const u32 reg = blockIdx.x, b = reg / in.subs;
  const u64 nb = min((u64)in.fill[(u64)reg * in.fill_stride], in.cap);
  const u64 base = (u64)blockIdx.y * kTileRecs;
  if (base >= nb) return;
  const u32 cnt = (u32)min((u64)kTileRecs, nb - base);
  const u64 off = (u64)reg * in.cap + base;
  const u32 fmask = (1u << g.fine_shift) - 1u;
  u64 h[R];
  low_t a[R];
  u32 dig[R];
  u32 valid = 0;
  load_tile<NT, R>(h, in.h + off, cnt);
  load_tile<NT, R>(a, in.a ? in.a + off : nullptr, cnt);
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const bool ok = threadIdx.x + NT * j < cnt;
    dig[j] = ok ? (u32)(seg_cell(h[j], p, g.x) >> g.slice_shift) & fmask : 0u;
    valid |= (ok ? 1u : 0u) << j;
  }
  tile_scatter<NT, R, 1 << kFineShift>(s, h, a, dig, valid, fmask + 1u, b << g.fine_shift, 0, out, side);
}

// Stage slice f of third p (its low cell bits, padded to kSliceCells) in
// LDS with 16-byte loads, all of a thread's in flight at once; returns the
// slice's length in real cells.  The block has NT threads.
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
template <int NT>
__device__ __forceinline__ u32 stage_slice(low_t* sl, const PartGeom& g, int p, u32 f) {
  constexpr int V = kSliceCells * sizeof(low_t) / 16 / NT;  // 16-byte words per thread
  static_assert(V * NT * 16 == kSliceCells * sizeof(low_t), "slice shape");
  const u64 c0 = (u64)f << g.slice_shift;
  const u64 len = c0 < g.x.block ? min((u64)kSliceCells, g.x.block - c0) : 0;
  if (len) {
    const u32x4v* src = (const u32x4v*)(g.lo + (u64)p * g.lo_third + c0);
    u32x4v v[V];
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = __builtin_nontemporal_load(src + threadIdx.x + NT * k);
#pragma unroll
    for (int k = 0; k < V; ++k) ((u32x4v*)sl)[threadIdx.x + NT * k] = v[k];
  }
  __syncthreads();
  return (u32)len;
}

// LDS of the lookup kernel's tile pipeline: one staging area, rankings of two
// tiles (the one being written out and the next)
template <int NB>
struct LookSmem {
  u64 h[kTileRecs];
  low_t a[kTileRecs];
  u16 d[kTileRecs];
  u32 hist[2][NB];
  u32 lofs[2][NB];
  u32 gbase[NB];
  u32 total[2];
};

// Exclusive scan of nbins bucket counts into lofs, their sum into *total;
// called by the block's first wavefront (threads 0-63) only
__device__ __forceinline__ void scan_bins(const u32* hist, u32* lofs, u32* total, u32 nbins) {
  const u32 t = threadIdx.x;
  const u32 per = (nbins + 63) / 64;
  u32 sum = 0;
  for (u32 k = 0; k < per; ++k) {
    const u32 b = t * per + k;
    if (b < nbins) sum += hist[b];
  }
  u32 incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const u32 v = (u32)__shfl_up((int)incl, o, 64);
    if ((int)t >= o) incl += v;
  }
  u32 run = incl - sum;
  for (u32 k = 0; k < per; ++k) {
    const u32 b = t * per + k;
    if (b < nbins) {
      lofs[b] = run;
      run += hist[b];
    }
  }
  if (t == 63) *total = incl;
}

// Each loaded record's cell of pass p (in slice [c0, c0 + len)) from LDS into
// its partial xor, then the coarse bucket of its next cell; returns the valid
// mask.  A record whose cell is not in the slice (not produced by the
// partition kernels) gets the full query instead.
template <int NT, int R>
__device__ __forceinline__ u32 look_tile(const PartGeom& g, int p, const low_t* sl, u64 c0, u32 len, u32 cnt,
                                         const u64 (&h)[R], low_t (&a)[R], u32 (&dig)[R], const PartSide& side) {
  u32 valid = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    dig[j] = 0;
    if (threadIdx.x + NT * j < cnt) {
      const u64 li = seg_cell(h[j], p, g.x) - c0;
      if (li < len) {
        a[j] ^= sl[li];
        dig[j] = (u32)(seg_cell(h[j], p + 1, g.x) >> (g.slice_shift + g.fine_shift));
        valid |= 1u << j;
      } else {
        side_push(side, h[j]);
      }
    }
  }
  return valid;
}

// One block per fine bucket (= slice) of pass p < 2: each record's cell of
// this pass from LDS into its partial xor, then on to the coarse bucket of
// its next cell.  The tiles of the bucket go through a software pipeline
// with two barriers per tile, so that no load or reservation round trip is
// waited for in full: while tile k is staged and written out, tile k+1 is
// looked up, ranked, scanned and reserved and tile k+2 is loaded.
template <int NB>
__global__ __launch_bounds__(kLookThreads) void k_part_lookup(int p, PartGeom g, PartSet in, PartSet out,
                                                               PartSide side) {
  constexpr int NT = kLookThreads, R = kTileRecs / NT;
  __shared__ LookSmem<NB> s;
  __shared__ __align__(16) low_t sl[kSliceCells];
  const u32 f = blockIdx.x;
  const u64 n = min((u64)in.fill[(u64)f * in.fill_stride], in.cap);
  if (n == 0) return;
  const u32 t = threadIdx.x, nbins = g.nc, sub = f % out.subs;
  // both count buffers start at zero (published by the slice's barrier)
  for (u32 b = t; b < nbins; b += NT) s.hist[0][b] = s.hist[1][b] = 0;
  const u32 len = stage_slice<NT>(sl, g, p, f);
  const u64 c0 = (u64)f << g.slice_shift;
  const u64 off0 = (u64)f * in.cap;
  const u64 ntiles = (n + kTileRecs - 1) / kTileRecs;
  auto cnt_of = [&](u64 k) { return (u32)min((u64)kTileRecs, n - k * kTileRecs); };
  auto reserve = [&](u32 b, u32 c) {
    return c ? atomicAdd(out.fill + ((u64)b * out.subs + sub) * out.fill_stride, c) : 0u;
  };
  // tile 0: loaded, looked up, ranked, scanned and reserved; tile 1 loading
  u64 hc[R];
  low_t ac[R];
  u32 dc[R], rc[R];
  load_tile<NT, R>(hc, in.h + off0, cnt_of(0));
  load_tile<NT, R>(ac, in.a ? in.a + off0 : nullptr, cnt_of(0));
  u32 vc = look_tile<NT, R>(g, p, sl, c0, len, cnt_of(0), hc, ac, dc, side);
#pragma unroll
  for (int j = 0; j < R; ++j) rc[j] = ((vc >> j) & 1u) ? atomicAdd(&s.hist[0][dc[j]], 1u) : 0u;
  __syncthreads();
  if (t < 64) scan_bins(s.hist[0], s.lofs[0], &s.total[0], nbins);
  u32 gb = t < nbins ? reserve(t, s.hist[0][t]) : 0u;
  __syncthreads();
  u64 hn[R];
  low_t an[R];
  if (ntiles > 1) {
    load_tile<NT, R>(hn, in.h + off0 + kTileRecs, cnt_of(1));
    load_tile<NT, R>(an, in.a ? in.a + off0 + kTileRecs : nullptr, cnt_of(1));
  }
  for (u64 k = 0; k < ntiles; ++k) {
    const u32 cur = (u32)(k & 1), nxt = cur ^ 1u;
    const bool more = k + 1 < ntiles;  // uniform over the block
    u64 hx[R];
    low_t ax[R];
    u32 dx[R], rx[R], vx = 0;
    if (more) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        hx[j] = hn[j];
        ax[j] = an[j];
      }
      vx = look_tile<NT, R>(g, p, sl, c0, len, cnt_of(k + 1), hx, ax, dx, side);
      if (k + 2 < ntiles) {
        const u64 o = off0 + (k + 2) * kTileRecs;
        load_tile<NT, R>(hn, in.h + o, cnt_of(k + 2));
        load_tile<NT, R>(an, in.a ? in.a + o : nullptr, cnt_of(k + 2));
      }
    }
    // tile k's counts were last read by the scan and the reservations of the
    // previous iteration (before its last barrier): zeroed here for tile k+2
    for (u32 b = t; b < nbins; b += NT) s.hist[cur][b] = 0;
    // tile k: reservation and staging; tile k+1: ranks (its count buffer was
    // zeroed in the previous iteration)
    if (t < nbins) s.gbase[t] = gb;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if ((vc >> j) & 1u) {
        const u32 pos = s.lofs[cur][dc[j]] + rc[j];
        s.h[pos] = hc[j];
        s.a[pos] = ac[j];
        s.d[pos] = (u16)dc[j];
      }
    if (more) {
#pragma unroll
      for (int j = 0; j < R; ++j) rx[j] = ((vx >> j) & 1u) ? atomicAdd(&s.hist[nxt][dx[j]], 1u) : 0u;
    }
    __syncthreads();
    // tile k+1: offsets and reservation (the offsets are used after the
    // iteration's last barrier, the reservation in the next iteration)
    if (more) {
      if (t < 64) scan_bins(s.hist[nxt], s.lofs[nxt], &s.total[nxt], nbins);
      if (t < nbins) gb = reserve(t, s.hist[nxt][t]);
    }
    // tile k: write-out in runs per bucket
    const u32 total = s.total[cur];
    for (u32 i = t; i < total; i += NT) {
      const u32 b = s.d[i];
      const u64 kk = (u64)s.gbase[b] + (i - s.lofs[cur][b]);
      const u64 at = ((u64)b * out.subs + sub) * out.cap + kk;
      if (kk < out.cap) {
        out.h[at] = s.h[i];
        if (out.a) out.a[at] = s.a[i];
      } else {
        side_push(side, s.h[i]);
      }
    }
    __syncthreads();  // staging, gbase and offsets are reused by the next iteration
    if (more) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        hc[j] = hx[j];
        ac[j] = ax[j];
        dc[j] = dx[j];
        rc[j] = rx[j];
      }
      vc = vx;
    }
  }
}

// One block per fine bucket of the third cell: a record whose partial xor
// matches the low fingerprint bits is a candidate for the full query.  The
// candidates (2^-8 of the records) gather in an LDS buffer, and the block
// reserves room in the side list once per flush: one global atomic per
// block and flush instead of one per candidate (a single counter serialises
// its atomics).  Held to 64 VGPRs (8 waves per SIMD), so that two blocks,
// each with its 64 KiB slice and 8 KiB of candidates, share a CU.
constexpr u32 kCandBuf = 1024;
__global__ __launch_bounds__(kLookThreads, 8) void k_part_last(PartGeom g, PartSet in, PartSide side) {
  constexpr int NT = kLookThreads, R = 8;
  __shared__ __align__(16) low_t sl[kSliceCells];
  __shared__ u64 cand[kCandBuf];
  __shared__ u32 ncand, cbase;
  const u32 f = blockIdx.x;
  const u64 n = min((u64)in.fill[(u64)f * in.fill_stride], in.cap);
  if (n == 0) return;
  if (threadIdx.x == 0) ncand = 0;
  const u32 len = stage_slice<NT>(sl, g, 2, f);  // its barrier also publishes ncand
  const u64 c0 = (u64)f << g.slice_shift;
  const u64 off0 = (u64)f * in.cap;
  // the buffered candidates to the side list (every thread calls it)
  auto flush = [&]() {
    __syncthreads();
    const u32 m = min(ncand, kCandBuf);
    if (threadIdx.x == 0) cbase = m ? atomicAdd(side.n, m) : 0u;
    __syncthreads();
    for (u32 i = threadIdx.x; i < m; i += NT) {
      if (cbase + i < side.cap)
        side.h[cbase + i] = cand[i];
      else
        atomicAdd(side.lost, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) ncand = 0;
    __syncthreads();
  };
  for (u64 base = 0; base < n; base += NT * R) {
    const u32 cnt = (u32)min((u64)NT * R, n - base);
    u64 h[R];
    low_t a[R];
    load_tile<NT, R>(h, in.h + off0 + base, cnt);
    load_tile<NT, R>(a, in.a + off0 + base, cnt);
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (threadIdx.x + NT * j < cnt) {
        const u64 li = seg_cell(h[j], 2, g.x) - c0;
        if (li >= len || (((u32)(a[j] ^ sl[li]) ^ (u32)h[j]) & g.mask_lo) == 0u) {
          const u32 k = atomicAdd(&ncand, 1u);  // LDS
          if (k < kCandBuf)
            cand[k] = h[j];
          else
            side_push(side, h[j]);  // more than the buffer in one round (small r): one atomic each
        }
      }
    // flush once the buffer is half full, so a round (NT * R records) cannot overrun it at r >= 8
    __syncthreads();
    if (ncand >= kCandBuf / 2) flush();
  }
  flush();
}

// full queries of the side list; the accepted hashes go to acc
__global__ void k_part_check(PartGeom g, PartSide side, PartSide acc) {
  const u32 n = min(*side.n, side.cap);
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const u64 h = side.h[i];
    if (query_hash(g.x, h)) side_push(acc, h);
  }
}

// h of k_probe_only's key i (salt): mix64(mix64(i * phi + salt) + seed)
__global__ void k_part_gen(XorDev x, u64 n, u64 salt, u64* __restrict__ h) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    h[i] = mix64(mix64(i * 0x9E3779B97F4A7C15ULL + salt) + x.seed);
}

// direct queries: cnt_sum[0] += accepted, cnt_sum[1] += sum of accepted hashes
__global__ __launch_bounds__(256) void k_part_ref(XorDev x, const u64* __restrict__ h, u64 n,
                                                  unsigned long long* cnt_sum) {
  u64 c = 0, sum = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u64 v = h[i];
    if (query_hash(x, v)) {
      ++c;
      sum += v;
    }
  }
  c = wave_sum(c);
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0 && c) {
    atomicAdd(cnt_sum, (unsigned long long)c);
    atomicAdd(cnt_sum + 1, (unsigned long long)sum);
  }
}

__global__ void k_part_sum(const u64* __restrict__ h, const u32* n, u32 cap, unsigned long long* cnt_sum) {
  const u32 m = min(*n, cap);
  u64 c = 0, sum = 0;
  for (u32 i = threadIdx.x; i < m; i += blockDim.x) {
    ++c;
    sum += h[i];
  }
  c = wave_sum(c);
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0 && c) {
    atomicAdd(cnt_sum, (unsigned long long)c);
    atomicAdd(cnt_sum + 1, (unsigned long long)sum);
  }
}

int launch_part_low(const PartGeom& g, low_t* lo, hipStream_t s) {
  if (g.lo_third < g.x.block || (g.lo_third & (kSliceCells - 1))) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_part_low, dim3(4096), dim3(256), 0, s, g.x.cells, g.x.block, g.lo_third, lo);
  return (int)hipGetLastError();
}

int launch_part_first(const u64* h, u64 n, const u32* tile_fill, const PartGeom& g, const PartSet& out,
                      const PartSide& side, hipStream_t s) {
  static_assert(kTileRecs == (int)kEmitChunk, "one emit chunk per partition tile");
  const u64 tiles = (n + kTileRecs - 1) / kTileRecs;
  if (tiles == 0) return 0;
  if (tiles > 0x7FFFFFFFull || g.nc == 0 || g.nc > kPartMaxBins || out.nb < g.nc || out.subs == 0)
    return (int)hipErrorInvalidValue;
  if (g.nc <= 256)
    hipLaunchKernelGGL(k_part_first<256>, dim3((u32)tiles), dim3(kPartThreads), 0, s, h, n, tile_fill, g, out, side);
  else
    hipLaunchKernelGGL(k_part_first<kPartMaxBins>, dim3((u32)tiles), dim3(kPartThreads), 0, s, h, n, tile_fill, g,
                       out, side);
  return (int)hipGetLastError();
}

int launch_part_fine(int pass, const PartGeom& g, const PartSet& in, const PartSet& out, const PartSide& side,
                     hipStream_t s) {
  const u64 tiles = (in.cap + kTileRecs - 1) / kTileRecs;
  const u64 regions = (u64)in.nb * in.subs;
  if (pass < 0 || pass > 2 || tiles == 0 || tiles > 65535 || in.nb != g.nc || regions == 0 ||
      regions > 0x7FFFFFFFull || out.nb < (g.nc << g.fine_shift) || out.subs != 1)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_part_fine, dim3((u32)regions, (u32)tiles), dim3(kPartThreads), 0, s, pass, g, in, out,
                     side);
  return (int)hipGetLastError();
}

int launch_part_lookup(int pass, const PartGeom& g, const PartSet& in, const PartSet& out, const PartSide& side,
                       hipStream_t s) {
  if (pass < 0 || pass > 1 || in.nb != (g.nc << g.fine_shift) || in.subs != 1 || out.nb < g.nc ||
      out.subs == 0 || g.nc > kPartMaxBins)
    return (int)hipErrorInvalidValue;
  if (g.nc <= 256)
    hipLaunchKernelGGL(k_part_lookup<256>, dim3(in.nb), dim3(kLookThreads), 0, s, pass, g, in, out, side);
  else
    hipLaunchKernelGGL(k_part_lookup<kPartMaxBins>, dim3(in.nb), dim3(kLookThreads), 0, s, pass, g, in, out, side);
  return (int)hipGetLastError();
}

int launch_part_last(const PartGeom& g, const PartSet& in, const PartSide& side, hipStream_t s) {
  if (in.nb != (g.nc << g.fine_shift) || in.subs != 1 || !in.a) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_part_last, dim3(in.nb), dim3(kLookThreads), 0, s, g, in, side);
  return (int)hipGetLastError();
}

int launch_part_check(const PartGeom& g, const PartSide& side, const PartSide& acc, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_part_check, dim3(grid), dim3(256), 0, s, g, side, acc);
  return (int)hipGetLastError();
}

int launch_part_gen(const XorDev& x, u64 n, u64 salt, u64* h, hipStream_t s) {
  hipLaunchKernelGGL(k_part_gen, dim3(8192), dim3(256), 0, s, x, n, salt, h);
  return (int)hipGetLastError();
}

int launch_part_ref(const XorDev& x, const u64* h, u64 n, unsigned long long* cnt_sum, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_part_ref, dim3(grid), dim3(256), 0, s, x, h, n, cnt_sum);
  return (int)hipGetLastError();
}

int launch_part_sum(const u64* h, const u32* n, u32 cap, unsigned long long* cnt_sum, hipStream_t s) {
  hipLaunchKernelGGL(k_part_sum, dim3(1), dim3(1024), 0, s, h, n, cap, cnt_sum);
  return (int)hipGetLastError();
}

}  // namespace fc5
