// fc5_exact.hip -- the exact sorted pair table on the device.
//
// The reference's exact index (QuadSumIndex, quad_sum.hpp:8-45) holds every
// pair sum a^6 + b^6 with 1 <= a <= b <= N sorted by sum, and equal_sum_range
// answers which pairs make one sum.  Here an entry is a single 128-bit word
// sum * 2^s + a (s = bit length of N): the sum orders the entries, a names the
// pair (b is the exact sixth root of sum - a^6), and a = 0 marks an extra value
// (a planted target).  Row a (b = a..N) is written coalesced at its closed-form
// offset, the array is radix-sorted on its significant bits only (this file's
// LSD radix sort, 8-bit digits, over a double buffer), and a lookup is one
// binary search per target.
//
// Radix sort, one pass per digit: k_sort_hist counts each 4096-key tile's
// digits (digit-major, counts[d * tiles + t]); k_sort_scan gives every
// (digit, tile) its offset within the digit (one block per digit) and the
// digit's total; k_sort_scatter sorts its tile by digit in LDS (stable: keys
// ranked in index order, wave-level digit match by ballots) and writes each
// digit's run to base[d] + offset[d][t] coalesced.
#include <algorithm>

#include "fc5_device.hpp"

namespace fc5 {

namespace {

// row a starts after rows 1..a-1 of N - a' + 1 entries each
__global__ __launch_bounds__(256) void k_exact_rows(u128* __restrict__ keys, u32 N, u32 s) {
  for (u32 a = blockIdx.x + 1; a <= N; a += gridDim.x) {
    const u64 am1 = a - 1u;
    const u64 off = am1 * ((u64)N + 1) - am1 * a / 2;
    const u128 pa = p6_u128(a);
    for (u32 b = a + threadIdx.x; b <= N; b += blockDim.x) keys[off + (b - a)] = ((pa + p6_u128(b)) << s) | a;
  }
}

// one thread per target: lower bound of T * 2^s, then the run of entries whose
// value is T; pairs with b > hi are left out, an extra value is reported as
// (job, 0, 0) with pad = 1
__global__ __launch_bounds__(256) void k_exact_lookup(const u128* __restrict__ keys, u64 n, u32 s,
                                                      const PairJob* __restrict__ jobs, u32 nj,
                                                      PairHit* __restrict__ out, u32 cap, u32* count) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const PairJob J = jobs[j];
  const u128 T = ((u128)J.T_hi << 64) | J.T_lo;
  if (T == 0 || (T >> (128u - s)) != 0) return;  // no entry can hold T
  const u128 want = T << s;
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  const u128 amask = ((u128)1 << s) - 1;
  for (u64 k = lo; k < n; ++k) {
    const u128 key = keys[k];
    if ((key >> s) != T) break;
    const u32 a = (u32)(key & amask);
    u32 b = 0;
    if (a) {
      b = iroot6_u128(T - p6_u128(a));
      if (b > J.hi) continue;
    }
    const u32 slot = atomicAdd(count, 1u);
    if (slot < cap) out[slot] = PairHit{j, a, b, a ? 0u : 1u};
  }
}

constexpr u32 kSortThreads = 256;
constexpr u32 kSortTile = 4096;  // keys per tile: 16 per thread, 64 KB of LDS while sorting
constexpr u32 kBins = 256;       // 8-bit digits
constexpr u32 kScanThreads = 1024;

__device__ __forceinline__ u32 digit_of(u128 k, u32 shift) { return (u32)(k >> shift) & (kBins - 1u); }

__global__ __launch_bounds__(256) void k_sort_hist(const u128* __restrict__ keys, u64 n, u32 shift,
                                                   u32* __restrict__ counts, u64 tiles) {
  __shared__ u32 h[kBins];
  const u64 t = blockIdx.x;
  h[threadIdx.x] = 0;
  __syncthreads();
  const u64 base = t * kSortTile;
  const u32 cnt = (u32)min((u64)kSortTile, n - base);
  for (u32 i = threadIdx.x; i < cnt; i += kSortThreads) atomicAdd(&h[digit_of(keys[base + i], shift)], 1u);
  __syncthreads();
  counts[(u64)threadIdx.x * tiles + t] = h[threadIdx.x];
}

// one block per digit: exclusive prefix of its tile counts (u64) and its total
__global__ __launch_bounds__(1024) void k_sort_scan(const u32* __restrict__ counts, u64* __restrict__ offs, u64 tiles,
                                                    u64* __restrict__ totals) {
  __shared__ u64 wsum[kScanThreads / 64];
  __shared__ u64 carry;
  const u32 d = blockIdx.x, lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const u32* c = counts + (u64)d * tiles;
  u64* o = offs + (u64)d * tiles;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (u64 c0 = 0; c0 < tiles; c0 += kScanThreads) {
    const u64 i = c0 + threadIdx.x;
    const u64 v = i < tiles ? c[i] : 0u;
    u64 x = v;  // inclusive scan within the wavefront
    for (int k = 1; k < 64; k <<= 1) {
      const u64 y = __shfl_up(x, k, 64);
      if (lane >= (u32)k) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    u64 before = carry;
    for (u32 k = 0; k < w; ++k) before += wsum[k];
    if (i < tiles) o[i] = before + x - v;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) carry = before + x;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[d] = carry;
}

__global__ __launch_bounds__(256) void k_sort_scatter(const u128* __restrict__ in, u128* __restrict__ out, u64 n,
                                                      u32 shift, const u32* __restrict__ counts,
                                                      const u64* __restrict__ offs, const u64* __restrict__ totals,
                                                      u64 tiles) {
  __shared__ u128 s_keys[kSortTile];
  __shared__ u32 s_start[kBins], s_run[kBins], s_wcnt[4][kBins];
  __shared__ u64 s_goff[kBins], s_wtot[4];
  const u64 t = blockIdx.x;
  const u32 tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const u64 base = t * kSortTile;
  const u32 cnt = (u32)min((u64)kSortTile, n - base);
  {  // the tile's digit starts (exclusive prefix of its counts), the digits' starts in the
     // whole array (exclusive prefix of the totals) and so each digit's global offset
    const u32 c = counts[(u64)tid * tiles + t];
    const u64 tot = totals[tid];
    u32 x = c;
    u64 y = tot;
    for (int k = 1; k < 64; k <<= 1) {
      const u32 xs = (u32)__shfl_up((int)x, k, 64);
      const u64 ys = __shfl_up(y, k, 64);
      if (lane >= (u32)k) {
        x += xs;
        y += ys;
      }
    }
    if (lane == 63) {
      s_wcnt[w][0] = x;  // wave totals (s_wcnt is scratch here)
      s_wtot[w] = y;
    }
    __syncthreads();
    u32 before = 0;
    u64 db = 0;
    for (u32 k = 0; k < w; ++k) {
      before += s_wcnt[k][0];
      db += s_wtot[k];
    }
    s_start[tid] = before + x - c;
    s_goff[tid] = db + (y - tot) + offs[(u64)tid * tiles + t];
    s_run[tid] = 0;
    __syncthreads();
    for (u32 k = 0; k < 4; ++k) s_wcnt[k][tid] = 0;
    __syncthreads();
  }
  const u64 below = lane ? (~0ULL >> (64u - lane)) : 0ULL;
  for (u32 r = 0; r < kSortTile / kSortThreads; ++r) {
    const u32 i = r * kSortThreads + tid;
    const bool valid = i < cnt;
    const u128 key = valid ? in[base + i] : (u128)0;
    const u32 d = digit_of(key, shift);
    u64 m = __ballot(valid);
    for (u32 b = 0; b < 8; ++b) {
      const u64 bb = __ballot(valid && ((d >> b) & 1u));
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const u32 rank = (u32)__popcll(m & below);
    if (valid && rank == 0) s_wcnt[w][d] = (u32)__popcll(m);
    __syncthreads();
    if (valid) {
      u32 pos = s_start[d] + s_run[d] + rank;
      for (u32 k = 0; k < w; ++k) pos += s_wcnt[k][d];
      s_keys[pos] = key;
    }
    __syncthreads();
    s_run[tid] += s_wcnt[0][tid] + s_wcnt[1][tid] + s_wcnt[2][tid] + s_wcnt[3][tid];
    for (u32 k = 0; k < 4; ++k) s_wcnt[k][tid] = 0;
    __syncthreads();
  }
  for (u32 p = tid; p < cnt; p += kSortThreads) {
    const u128 key = s_keys[p];
    const u32 d = digit_of(key, shift);
    out[s_goff[d] + (p - s_start[d])] = key;
  }
}

}  // namespace

int launch_exact_rows(u128* keys, u32 N, u32 s, hipStream_t st) {
  const u32 blocks = N < 65536u ? N : 65536u;
  hipLaunchKernelGGL(k_exact_rows, dim3(blocks ? blocks : 1), dim3(256), 0, st, keys, N, s);
  return (int)hipGetLastError();
}

int exact_sort(u128* keys, u128* alt, u64 n, u32 end_bit, void* temp, size_t& temp_bytes, hipStream_t st,
               bool& sorted_in_alt) {
  const u64 tiles = (n + kSortTile - 1) / kSortTile;
  const size_t need = (size_t)tiles * kBins * (sizeof(u32) + sizeof(u64)) + kBins * sizeof(u64);
  if (!temp) {
    temp_bytes = std::max<size_t>(need, 16);
    return (int)hipSuccess;
  }
  if (temp_bytes < need || end_bit > 128 || tiles > 0xFFFFFFFFull) return (int)hipErrorInvalidValue;
  sorted_in_alt = false;
  if (!n) return (int)hipSuccess;
  u64* offs = static_cast<u64*>(temp);
  u64* totals = offs + tiles * kBins;
  u32* counts = reinterpret_cast<u32*>(totals + kBins);
  u128 *src = keys, *dst = alt;
  for (u32 shift = 0; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_sort_hist, dim3((u32)tiles), dim3(kSortThreads), 0, st, src, n, shift, counts, tiles);
    hipLaunchKernelGGL(k_sort_scan, dim3(kBins), dim3(kScanThreads), 0, st, counts, offs, tiles, totals);
    hipLaunchKernelGGL(k_sort_scatter, dim3((u32)tiles), dim3(kSortThreads), 0, st, src, dst, n, shift, counts, offs,
                       totals, tiles);
    std::swap(src, dst);
    sorted_in_alt = !sorted_in_alt;
  }
  return (int)hipGetLastError();
}

int launch_exact_lookup(const u128* keys, u64 n, u32 s, const PairJob* jobs, u32 nj, PairHit* out, u32 cap,
                        u32* count, hipStream_t st) {
  const u32 blocks = nj ? (nj + 255) / 256 : 1;
  hipLaunchKernelGGL(k_exact_lookup, dim3(blocks), dim3(256), 0, st, keys, n, s, jobs, nj, out, cap, count);
  return (int)hipGetLastError();
}

}  // namespace fc5
