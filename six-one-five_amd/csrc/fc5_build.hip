// fc5_build.hip -- xor pair table built on the device (gfx950).
//
// Same key set, capacity rule, hashing and seed schedule as the host builder
// (spike/ribbon/xor_filter.hpp:185-192,256-470; fourcore_xor_store.hpp:67-119),
// so the result is a valid xor filter over exactly the same keys; its bytes
// differ from the host builder's because the peel order differs.
//
//   1. keys a^6+b^6 (mod 2^64), 1<=a<=b<=N, generated in registers (one
//      workgroup per row a) and deduplicated in an open-addressing hash set
//      of 1.25 slots per key (atomicCAS, linear probing from a multiply-shift
//      position; key 0 kept in a flag).  Duplicates are systematic:
//      a^6 = 0 (mod 2^64) whenever 2^11 | a.
//   2. incidence: per unique key, atomic xor of its hash and atomic count on
//      its three cells.  Counts are 16-bit, two per 32-bit word (a count
//      reaching 2^16 fails the build, as the reference's u16 counts do).
//   3. peel in rounds.  claim: every cell of the frontier with count 1 reads
//      its key and keeps it iff it is the key's first cell with count 1 (the
//      counts are not written during a claim pass, so each key is claimed
//      once); apply: the claimed keys are removed from their other cells,
//      whose counts dropping to 1 form the next frontier, written over the
//      one just claimed (one frontier buffer of m/2 cells; a larger frontier
//      fails the build).  The claimed cell keeps the key's hash.
//
// Device working set: 8 B x 1.25 per key (hash set) + 8 B per key (peel
// order) + 8 + 2 + 4 B per cell (xor sets, counts, frontier) -- about 25 B per
// key, 250 GB at N = 119,047 (7.09e9 keys).
//   4. assign in reverse round order, in place: cell = fp(h) ^ other cells.
//      Keys claimed in the same round never share cells, so a round is one
//      parallel pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "fc5_build.hpp"
#include "fc5_mix.hpp"
#include "fc5_moddiv.hpp"

namespace fc5 {

typedef unsigned long long ull;

__device__ __forceinline__ u64 shfl64(u64 v, int lane) {
  const u32 lo = (u32)__shfl((int)(u32)v, lane, 64), hi = (u32)__shfl((int)(u32)(v >> 32), lane, 64);
  return ((u64)hi << 32) | lo;
}

// wave-aggregated append: returns this lane's slot (meaningful when pred)
__device__ __forceinline__ u64 wave_append(ull* ctr, bool pred) {
  const u64 mask = __ballot(pred);
  if (!mask) return 0;
  const int leader = __ffsll((long long)mask) - 1;
  const int lane = (int)(threadIdx.x & 63u);
  u64 base = 0;
  if (lane == leader) base = atomicAdd(ctr, (ull)__popcll(mask));
  base = shfl64(base, leader);
  return base + __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

__global__ void k_p6(u64* p6, u32 N) {
  for (u32 x = blockIdx.x * blockDim.x + threadIdx.x; x <= N; x += gridDim.x * blockDim.x) {
    const u64 x2 = (u64)x * x;
    p6[x] = x2 * x2 * x2;
  }
}

// the set holds more slots than keys, so every probe sequence ends
__device__ __forceinline__ void set_insert(u64* slots, u64 nslots, u32* zero_flag, u64 key) {
  if (key == 0) {
    atomicOr(zero_flag, 1u);
    return;
  }
  u64 pos = __umul64hi(mix64(key ^ 0x2545F4914F6CDD1DULL), nslots);
  for (;;) {
    const u64 prev = atomicCAS((ull*)&slots[pos], 0ULL, (ull)key);
    if (prev == 0 || prev == key) return;
    if (++pos == nslots) pos = 0;
  }
}

__global__ void k_insert_pairs(const u64* p6, u32 N, u64* slots, u64 nslots, u32* zero_flag) {
  for (u32 a = blockIdx.x + 1; a <= N; a += gridDim.x) {
    const u64 pa = p6[a];
    for (u32 b = a + threadIdx.x; b <= N; b += blockDim.x) set_insert(slots, nslots, zero_flag, pa + p6[b]);
  }
}

__global__ void k_insert_keys(const u64* keys, u64 n, u64* slots, u64 nslots, u32* zero_flag) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    set_insert(slots, nslots, zero_flag, keys[i]);
}

// 16-bit cell counts, two per 32-bit word
__device__ __forceinline__ u32 cnt_get(const u32* counts, u64 i) {
  return (counts[i >> 1] >> ((u32)(i & 1) * 16u)) & 0xffffu;
}
__device__ __forceinline__ u32 cnt_add(u32* counts, u64 i, int delta) {  // returns the old count
  const u32 sh = (u32)(i & 1) * 16u;
  const u32 old = delta > 0 ? atomicAdd(&counts[i >> 1], 1u << sh) : atomicSub(&counts[i >> 1], 1u << sh);
  return (old >> sh) & 0xffffu;
}

__global__ void k_count_slots(const u64* slots, u64 n, ull* count) {
  u64 local = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    local += slots[i] != 0;
  for (int o = 32; o > 0; o >>= 1) local += shfl64(local, (int)((threadIdx.x & 63u) ^ (u32)o));
  if ((threadIdx.x & 63u) == 0) atomicAdd(count, (ull)local);
}

__device__ __forceinline__ void add_key(u64 key, u64 seed, u64 block, const ModDiv& md, u64* setxor, u32* counts,
                                        u32* overflow) {
  const u64 h = mix64(key + seed);
  u64 c[3];
  xor_cells(h, block, md, c);
  for (int t = 0; t < 3; ++t) {
    atomicXor((ull*)&setxor[c[t]], (ull)h);
    if (cnt_add(counts, c[t], 1) >= 0xffffu) atomicOr(overflow, 1u);
  }
}

__global__ void k_incidence(const u64* slots, u64 nslots, const u32* zero_flag, u64 seed, u64 block, ModDiv md,
                            u64* setxor, u32* counts, u32* overflow) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid == 0 && *zero_flag) add_key(0, seed, block, md, setxor, counts, overflow);
  for (u64 i = tid; i < nslots; i += (u64)gridDim.x * blockDim.x) {
    const u64 key = slots[i];
    if (key) add_key(key, seed, block, md, setxor, counts, overflow);
  }
}

__global__ void k_frontier_init(const u32* counts, u64 m, u64* flog, ull* flog_n, u64 flog_cap) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  // every lane of a wave runs the same trip count so the wave stays converged
  const u64 trips = (m + stride - 1) / stride;
  for (u64 t = 0; t < trips; ++t) {
    const u64 i = t * stride + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const bool one = i < m && cnt_get(counts, i) == 1;
    const u64 slot = wave_append(flog_n, one);
    if (one && slot < flog_cap) flog[slot] = i;
  }
}

__global__ void k_claim(const u64* front, u64 nf, const u32* counts, const u64* setxor, u64 block, ModDiv md,
                        u64* peel, ull* peel_n, u64 peel_cap) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  const u64 trips = (nf + stride - 1) / stride;
  for (u64 t = 0; t < trips; ++t) {
    const u64 k = t * stride + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    u64 i = 0;
    if (k < nf) {
      i = front[k];
      if (cnt_get(counts, i) == 1) {
        u64 c[3];
        xor_cells(setxor[i], block, md, c);
        const u64 first = cnt_get(counts, c[0]) == 1 ? c[0] : (cnt_get(counts, c[1]) == 1 ? c[1] : c[2]);
        take = first == i;
      }
    }
    const u64 slot = wave_append(peel_n, take);
    if (take && slot < peel_cap) peel[slot] = i;
  }
}

__global__ void k_apply(const u64* peel, u64 np, u32* counts, u64* setxor, u64 block, ModDiv md, u64* flog,
                        ull* flog_n, u64 flog_cap) {
  const u64 stride = (u64)gridDim.x * blockDim.x;
  const u64 trips = (np + stride - 1) / stride;
  for (u64 t = 0; t < trips; ++t) {
    const u64 k = t * stride + (u64)blockIdx.x * blockDim.x + threadIdx.x;
    bool nf[3] = {false, false, false};
    u64 c[3] = {0, 0, 0};
    if (k < np) {
      const u64 i = peel[k];
      const u64 h = setxor[i];
      xor_cells(h, block, md, c);
      for (int q = 0; q < 3; ++q) {
        if (c[q] == i) continue;
        atomicXor((ull*)&setxor[c[q]], (ull)h);
        nf[q] = cnt_add(counts, c[q], -1) == 2u;
      }
      (void)cnt_add(counts, i, -1);  // the claimed cell held this key only: 1 -> 0
    }
    for (int q = 0; q < 3; ++q) {
      const u64 slot = wave_append(flog_n, nf[q]);
      if (nf[q] && slot < flog_cap) flog[slot] = c[q];
    }
  }
}

__global__ void k_assign(const u64* peel, u64 np, u64* cells, u64 block, ModDiv md, u64 mask) {
  for (u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (u64)gridDim.x * blockDim.x) {
    const u64 i = peel[k];
    const u64 h = cells[i];
    u64 c[3];
    xor_cells(h, block, md, c);
    u64 v = h & mask;
    for (int q = 0; q < 3; ++q)
      if (c[q] != i) v ^= cells[c[q]];
    cells[i] = v & mask;
  }
}

// r-bit packing of expanded cells, 8 cells (r bytes) per thread
__global__ void k_pack(const u64* cells, u64 m, u32 r, u8* out) {
  const u64 groups = (m + 7) / 8;
  for (u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += (u64)gridDim.x * blockDim.x) {
    u64 buf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (u32 k = 0; k < 8; ++k) {
      const u64 idx = g * 8 + k;
      const u64 v = idx < m ? cells[idx] : 0;
      const u32 bit = k * r, w = bit >> 6, off = bit & 63u;
      buf[w] |= v << off;
      if (off && off + r > 64u) buf[w + 1] |= v >> (64u - off);
    }
    u8* dst = out + g * r;
    for (u32 j = 0; j < r; ++j) dst[j] = (u8)(buf[j >> 3] >> (8 * (j & 7)));
  }
}

static int grid_of(u64 n, int threads = 256) {
  const u64 g = (n + (u64)threads - 1) / (u64)threads;
  return (int)std::max<u64>(1, std::min<u64>(g, 32768));
}

static size_t capacity_rule(size_t n) {  // xor_filter.hpp:185-192
  const double need = 32.0 + 1.23 * (double)n;
  size_t cap = (size_t)(need + 0.999);
  if (cap < 3) cap = 3;
  return (cap + 2) / 3 * 3;
}

#define BCHK(expr)                                                                  \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr;       \
      goto fail;                                                                    \
    }                                                                               \
  } while (0)

int gpu_build_pairs(int N, u32 r, const u64* extra, size_t n_extra, hipStream_t s, GpuBuilt& out,
                    std::string& err) {
  out = GpuBuilt{};
  u64 *p6 = nullptr, *slots = nullptr, *d_extra = nullptr, *setxor = nullptr, *flog = nullptr, *peel = nullptr;
  u32 *counts = nullptr, *flags = nullptr;  // flags[0] zero key present, flags[1] count overflow
  ull* ctr = nullptr;                       // [0] unique, [1] flog_n, [2] peel_n
  const u64 mask = r >= 64 ? ~0ULL : ((1ULL << r) - 1);
  const u64 npairs = (u64)N * (u64)(N + 1) / 2;
  // 1.25 slots per key (rounded up to 64): linear probing stays short
  const u64 nslots = std::max<u64>(1024, ((npairs + n_extra) + (npairs + n_extra) / 4 + 63) / 64 * 64);
  u64 n_unique = 0, m = 0, block = 0, cwords = 0, fcap = 0;
  ModDiv md{};
  bool ok = false;
  ull h_ctr[3];
  u32 h_flags[2];

  BCHK(hipMalloc(&p6, ((u64)N + 1) * 8));
  BCHK(hipMalloc(&slots, nslots * 8));
  BCHK(hipMalloc(&flags, 8));
  BCHK(hipMalloc(&ctr, 3 * sizeof(ull)));
  BCHK(hipMemsetAsync(slots, 0, nslots * 8, s));
  BCHK(hipMemsetAsync(flags, 0, 8, s));
  BCHK(hipMemsetAsync(ctr, 0, 3 * sizeof(ull), s));
  hipLaunchKernelGGL(k_p6, dim3(grid_of((u64)N + 1)), dim3(256), 0, s, p6, (u32)N);
  hipLaunchKernelGGL(k_insert_pairs, dim3(std::min(N, 65536)), dim3(256), 0, s, p6, (u32)N, slots, nslots, flags);
  if (n_extra) {
    BCHK(hipMalloc(&d_extra, n_extra * 8));
    BCHK(hipMemcpyAsync(d_extra, extra, n_extra * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_insert_keys, dim3(grid_of(n_extra)), dim3(256), 0, s, d_extra, (u64)n_extra, slots,
                       nslots, flags);
  }
  hipLaunchKernelGGL(k_count_slots, dim3(grid_of(nslots)), dim3(256), 0, s, slots, nslots, ctr);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(h_ctr, ctr, sizeof(ull), hipMemcpyDeviceToHost, s));
  BCHK(hipMemcpyAsync(h_flags, flags, 8, hipMemcpyDeviceToHost, s));
  BCHK(hipStreamSynchronize(s));
  (void)hipFree(p6);
  p6 = nullptr;
  if (d_extra) { (void)hipFree(d_extra); d_extra = nullptr; }
  n_unique = h_ctr[0] + (h_flags[0] ? 1 : 0);
  m = capacity_rule(n_unique);
  block = m / 3;
  md = make_moddiv(block);
  cwords = (m + 1) / 2;             // two 16-bit counts per word
  fcap = std::max<u64>(m / 2, 1024);  // one frontier at a time
  BCHK(hipMalloc(&setxor, m * 8));
  BCHK(hipMalloc(&counts, cwords * 4));
  BCHK(hipMalloc(&flog, fcap * 8));
  BCHK(hipMalloc(&peel, std::max<u64>(n_unique, 1) * 8));
  for (int outer = 0; outer < 4 && !ok; ++outer) {
    const u64 seed = 0x615615615615615ULL + (u64)outer * 0x9E3779B97F4A7C15ULL;
    BCHK(hipMemsetAsync(setxor, 0, m * 8, s));
    BCHK(hipMemsetAsync(counts, 0, cwords * 4, s));
    BCHK(hipMemsetAsync(ctr, 0, 3 * sizeof(ull), s));
    BCHK(hipMemsetAsync(flags + 1, 0, 4, s));
    hipLaunchKernelGGL(k_incidence, dim3(grid_of(nslots)), dim3(256), 0, s, slots, nslots, flags, seed, block, md,
                       setxor, counts, flags + 1);
    hipLaunchKernelGGL(k_frontier_init, dim3(grid_of(m)), dim3(256), 0, s, counts, m, flog, ctr + 1, fcap);
    BCHK(hipGetLastError());
    std::vector<u64> round_start;
    u64 f_n = 0, p_done = 0;
    BCHK(hipMemcpyAsync(h_ctr, ctr, 3 * sizeof(ull), hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    f_n = h_ctr[1];
    if (f_n > fcap) { err = "frontier overflow"; goto fail; }
    while (f_n) {
      hipLaunchKernelGGL(k_claim, dim3(grid_of(f_n)), dim3(256), 0, s, flog, f_n, counts, setxor, block, md, peel,
                         ctr + 2, n_unique);
      // the claimed frontier is dead: the next one is written over it
      BCHK(hipMemsetAsync(ctr + 1, 0, sizeof(ull), s));
      BCHK(hipMemcpyAsync(h_ctr, ctr, 3 * sizeof(ull), hipMemcpyDeviceToHost, s));
      BCHK(hipStreamSynchronize(s));
      const u64 p_end = h_ctr[2];
      if (p_end > n_unique) { err = "peel list overflow"; goto fail; }
      round_start.push_back(p_done);
      if (p_end > p_done)
        hipLaunchKernelGGL(k_apply, dim3(grid_of(p_end - p_done)), dim3(256), 0, s, peel + p_done, p_end - p_done,
                           counts, setxor, block, md, flog, ctr + 1, fcap);
      BCHK(hipGetLastError());
      BCHK(hipMemcpyAsync(h_ctr, ctr, 3 * sizeof(ull), hipMemcpyDeviceToHost, s));
      BCHK(hipStreamSynchronize(s));
      p_done = p_end;
      f_n = h_ctr[1];
      if (f_n > fcap) { err = "frontier overflow"; goto fail; }
    }
    BCHK(hipMemcpyAsync(h_flags, flags, 8, hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    if (h_flags[1]) {
      err = "cell incidence count overflowed 16 bits";
      goto fail;
    }
    if (p_done != n_unique) continue;  // stalled: next seed base
    round_start.push_back(p_done);
    for (size_t rd = round_start.size() - 1; rd-- > 0;) {
      const u64 a = round_start[rd], b = round_start[rd + 1];
      if (b > a)
        hipLaunchKernelGGL(k_assign, dim3(grid_of(b - a)), dim3(256), 0, s, peel + a, b - a, setxor, block, md,
                           mask);
    }
    BCHK(hipGetLastError());
    BCHK(hipStreamSynchronize(s));
    out.rounds = round_start.size() - 1;
    out.hdr = StoreHeaderPod{};
    out.hdr.magic = 0x53353136u;
    out.hdr.version = 2;
    out.hdr.kind = 1;
    out.hdr.r = r;
    out.hdr.N = (u32)N;
    out.hdr.n_keys = n_unique;
    out.hdr.m_cells = m;
    out.hdr.mix_seed = seed;
    out.hdr.shard_count = 1;
    ok = true;
  }
  if (!ok) {
    err = "device xor construction stalled at every seed";
    goto fail;
  }
  (void)hipFree(slots);
  (void)hipFree(counts);
  (void)hipFree(flog);
  (void)hipFree(peel);
  (void)hipFree(flags);
  (void)hipFree(ctr);
  out.cells = setxor;
  return 0;
fail:
  for (void* p : {(void*)p6, (void*)slots, (void*)d_extra, (void*)setxor, (void*)flog, (void*)peel, (void*)counts,
                  (void*)flags, (void*)ctr})
    if (p) (void)hipFree(p);
  return -1;
}

int gpu_pack_cells(const u64* cells, u64 m, u32 r, u8* out_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3(grid_of((m + 7) / 8)), dim3(256), 0, s, cells, m, r, out_dev);
  return (int)hipGetLastError();
}

}  // namespace fc5
