// fc5_stage1.hip -- Stage-1 unit emission on the device (classes 2-5).
//
// The units of class c at B are the u in the u-band with u = s (mod M_c) for
// a seed s and the class's divisibility rules (fourcore_classes_v3.hpp:84-154,
// fourcore_hunt_v3.cpp:371-381).  The seeds are the sixth roots of B^6, i.e.
// B times the sixth roots of unity: mod 2^6 (4 odd roots), mod 3^6 (6) and
// mod 7^6 (6), combined by the Chinese remainder theorem (class 2: mod 14^6,
// class 3: mod 21^6, classes 4 and 5: mod 7^6).  The divisibility rules become
// one more congruence, so every class is a union of arithmetic progressions:
//   class 2  u = s (14^6), u = 3 or 6 (9)      -> 48 progressions of step 9*14^6
//   class 3  u = s (21^6), u even              -> 36 progressions of step 2*21^6
//   class 4/5 u = s (7^6), u = 0 (6)           ->  6 progressions of step 6*7^6
// One thread per B walks them in increasing order (each unit is the least
// progression member above the previous one), counting in the first pass and
// writing at its exclusive-scan offset in the second, so the output is the
// host emitter's list: B ascending, classes ascending, u ascending.
// The exclusive scan between the passes is this file's own: per block of
// 2048 counts a wave-shuffle scan of each lane's 8-count run, the block
// totals scanned by one block, then added back.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fc5_kernels.hpp"

namespace fc5 {

namespace {

constexpr u64 kM7 = 117649, kM14 = 7529536, kM21 = 85766121;

__device__ __forceinline__ u64 mulmod(u64 a, u64 b, u64 m) { return (u64)((unsigned __int128)a * b % m); }

// member of progression p (0 <= p < n_prog) of class c for B, and its step
__device__ __forceinline__ void progression(const Stage1Args& a, u64 B, int c, u32 p, u64& r, u64& step) {
  if (c == 2) {  // p = (j * 6 + l) * 2 + t
    const u32 t = p & 1u, jl = p >> 1, j = jl / 6u, l = jl % 6u;
    const u64 s2 = mulmod(B % 64u, a.z2[j], 64u), s7 = mulmod(B % kM7, a.z7[l], kM7);
    const u64 s = (s2 * a.e2_14 + s7 * a.e7_14) % kM14;
    const u64 want = t ? 6u : 3u;  // u mod 9; 14^6 = 1 (mod 9)
    r = s + ((want + 9u - s % 9u) % 9u) * kM14;
    step = 9u * kM14;
  } else if (c == 3) {  // p = k * 6 + l
    const u32 k = p / 6u, l = p % 6u;
    const u64 s3 = mulmod(B % 729u, a.z3[k], 729u), s7 = mulmod(B % kM7, a.z7[l], kM7);
    const u64 s = (s3 * a.e3_21 + s7 * a.e7_21) % kM21;
    r = s + (s & 1u) * kM21;  // even u; 21^6 is odd
    step = 2u * kM21;
  } else {  // classes 4 and 5: p = l
    const u64 s = mulmod(B % kM7, a.z7[p], kM7);
    r = s + ((6u - s % 6u) % 6u) * kM7;  // u = 0 (mod 6); 7^6 = 1 (mod 6)
    step = 6u * kM7;
  }
}

__device__ __forceinline__ u32 n_prog(int c) { return c == 2 ? 48u : c == 3 ? 36u : 6u; }

__global__ __launch_bounds__(256) void k_stage1(Stage1Args a, int pass) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  const u64 B = a.lo + i;
  u32 k = 0;
  const u64 base = pass ? a.offsets[i] : 0;
  if (B > 1 && (B & 1) && B % 3 && B % 7) {
    long long lo = (long long)(a.lo_frac * (double)B), hi = (long long)(a.hi_frac * (double)B);
    if (lo < 1) lo = 1;
    if (hi > (long long)B - 1) hi = (long long)B - 1;
    for (int c = 2; c <= 5; ++c) {
      if (!(a.cls_mask & (1u << c)) || lo > hi) continue;
      const u32 np = n_prog(c);
      u64 last = (u64)lo - 1;
      for (;;) {
        u64 best = ~0ull;
        for (u32 p = 0; p < np; ++p) {
          u64 r, step;
          progression(a, B, c, p, r, step);
          const u64 m = r > last ? r : r + ((last - r) / step + 1) * step;
          if (m <= (u64)hi && m < best) best = m;
        }
        if (best == ~0ull) break;
        if (pass) {
          a.out_cls[base + k] = (u32)c;
          a.out_B[base + k] = B;
          a.out_u[base + k] = best;
        }
        ++k;
        last = best;
      }
    }
  }
  if (!pass) a.counts[i] = k;
}

}  // namespace

int launch_stage1(const Stage1Args& a, int pass, hipStream_t s) {
  const u64 blocks = (a.count + 255) / 256;
  hipLaunchKernelGGL(k_stage1, dim3((u32)blocks), dim3(256), 0, s, a, pass);
  return (int)hipGetLastError();
}

namespace {

constexpr u32 kScanThreads = 256, kScanPer = 8, kScanBlock = kScanThreads * kScanPer;  // 2048 counts

// exclusive prefix of v over the block's threads (wave shuffles, then the
// four wave totals through LDS); *total gets the block's sum
__device__ __forceinline__ u64 block_exclusive(u64 v, u64* s_wave, u64* total) {
  const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  u64 x = v;
  for (u32 o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  u64 before = 0, all = 0;
  for (u32 k = 0; k < kScanThreads / 64; ++k) {
    if (k < w) before += s_wave[k];
    all += s_wave[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// block b: exclusive offsets of counts [2048 b, 2048 (b + 1)) within the block, its total to sums[b]
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const u64* in, u64* out, u64 n, u64* sums) {
  __shared__ u64 s_wave[kScanThreads / 64];
  const u64 base = (u64)blockIdx.x * kScanBlock + (u64)threadIdx.x * kScanPer;
  u64 v[kScanPer], run = 0;
#pragma unroll
  for (u32 k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    run += v[k];
  }
  u64 total;
  u64 off = block_exclusive(run, s_wave, &total);
#pragma unroll
  for (u32 k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// one block: exclusive scan of the m <= 2048 block totals, in place
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(u64* sums, u32 m) {
  __shared__ u64 s_wave[kScanThreads / 64];
  const u32 base = threadIdx.x * kScanPer;
  u64 v[kScanPer], run = 0;
#pragma unroll
  for (u32 k = 0; k < kScanPer; ++k) {
    v[k] = base + k < m ? sums[base + k] : 0;
    run += v[k];
  }
  u64 total;
  u64 off = block_exclusive(run, s_wave, &total);
#pragma unroll
  for (u32 k = 0; k < kScanPer; ++k) {
    if (base + k < m) sums[base + k] = off;
    off += v[k];
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(u64* out, u64 n, const u64* sums) {
  const u64 add = sums[blockIdx.x];
  const u64 base = (u64)blockIdx.x * kScanBlock;
  for (u32 k = threadIdx.x; k < kScanBlock; k += kScanThreads)
    if (base + k < n) out[base + k] += add;
}

}  // namespace

// exclusive prefix sum of n u64 counts: temp holds one u64 per 2048 counts,
// n up to 2048 * 2048 (temp == nullptr: temp_bytes gets the size)
int stage1_scan(const u64* counts_in, u64* offsets_out, u64 n, void* temp, size_t& temp_bytes, hipStream_t s) {
  const u64 blocks = (n + kScanBlock - 1) / kScanBlock;
  if (blocks > kScanBlock) return (int)hipErrorInvalidValue;
  if (!temp) {
    temp_bytes = (size_t)std::max<u64>(blocks, 1) * 8;
    return (int)hipSuccess;
  }
  if (!n) return (int)hipSuccess;
  if (temp_bytes < blocks * 8) return (int)hipErrorInvalidValue;
  u64* sums = static_cast<u64*>(temp);
  hipLaunchKernelGGL(k_scan_blocks, dim3((u32)blocks), dim3(kScanThreads), 0, s, counts_in, offsets_out, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, sums, (u32)blocks);
  hipLaunchKernelGGL(k_scan_add, dim3((u32)blocks), dim3(kScanThreads), 0, s, offsets_out, n, sums);
  return (int)hipGetLastError();
}

}  // namespace fc5
