// fc5_twin.hip -- the (6,1,7) class-5 twin search on gfx950 (MI355X).
//
// k_twin_batched is a persistent kernel over a device work queue; one work
// item is one (candidate, d) cell of the reference's grid
// (k_find4_cls5, solve_617_v1.cu:563-657).  For its T it takes 64 values of
// c4 at a time, one per lane: triple-sum gate on T - c4^6 (bitmaps mod 504
// and 247 in LDS), the exact c3 window of T - c4^6 in 128 bits; then, one c4
// after another, the 64 lanes sweep c3 through the pair-sum gate (LDS bitmap) and compact the
// survivors into a per-wave LDS ring that is probed 64 keys at a time
// (key = T - c4^6 - c3^6 mod 2^64, 3 gathers each).  Every c4 of a window is
// visited once.
#include "fc5_device.hpp"
#include "fc5_twin.hpp"

namespace fc5 {

namespace {

constexpr int kTwWaves = 4;
constexpr u32 kTwRing = 128;

struct TwRing {
  u32 *qc, *qd, *qx;  // candidate index, d, c3 | c4 << 16
  u32 head, tail;
};

__device__ __forceinline__ void tw_drain(const TwinArgs& a, const TwRing& q, u32 n, u32 lane) {
  if (lane >= n) return;
  const u32 slot = (q.head + lane) & (kTwRing - 1);
  const u32 ci = q.qc[slot], d = q.qd[slot], x = q.qx[slot];
  const u32 c3 = x & 0xffffu, c4 = x >> 16;
  const DevTwinCand* C = a.cands + ci;
  const u64 d3 = (u64)d * d * d;
  const u64 dq = (u64)(((u128)d3 * d3) >> 6);
  const u64 a2 = (u64)c4 * c4, b2 = (u64)c3 * c3;
  const u64 fp = C->Q_lo - dq - a2 * a2 * a2 - b2 * b2 * b2;
  if (d_query(a.main, fp)) {
    atomicAdd((unsigned long long*)&a.ctr->maybe, 1ULL);
    const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
    if (s < a.hit_cap) {
      TwinHit h;
      h.cand = C->uid;
      h.d = d;
      h.c4 = c4;
      h.c3 = c3;
      h.fp = fp;
      a.hits[s] = h;
    }
  }
}

__device__ __forceinline__ void tw_push(const TwinArgs& a, TwRing& q, bool pass, u32 ci, u32 d, u32 x,
                                        u32 lane) {
  const u64 m = __ballot(pass);
  if (!m) return;
  if (pass) {
    const u32 slot = (q.tail + lanes_below(m)) & (kTwRing - 1);
    q.qc[slot] = ci;
    q.qd[slot] = d;
    q.qx[slot] = x;
  }
  q.tail += (u32)__popcll(m);
  if (q.tail - q.head >= 64u) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    tw_drain(a, q, 64u, lane);
    q.head += 64u;
  }
}

__device__ __forceinline__ void twin_item(const TwinArgs& a, const u32* s_gate, const u32* s_tri, TwRing& q,
                                          u32 item, u32 lane) {
  u32 lo = 0, hi = a.n_cands;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (a.cands[mid].item_base <= item) lo = mid;
    else hi = mid;
  }
  const u32 ci = lo;
  const DevTwinCand* Cp = a.cands + ci;
  const u32 r = item - Cp->item_base;
  const u32 nres = Cp->nres, lim = Cp->lim;
  const u32 d = (u32)Cp->res[r % nres] + (r / nres) * 64u;
  if (d < 1 || d > Cp->fmax1) return;
  const u64 d3 = (u64)d * d * d;
  const u128 dq = ((u128)d3 * d3) >> 6;
  const u128 Q = ((u128)Cp->Q_hi << 64) | Cp->Q_lo;
  if (Q <= dq) return;  // R = base - (21d)^6 not positive
  const u128 T = Q - dq;
  const u32 hi4 = min(iroot6_u128(T), lim);
  if (hi4 < 1) return;
  const u32 lo4 = min_core(T, a.c4_div);
  if (lo4 > hi4) return;
  const u32 xg = mod_gate(T);
  u64 n_c4 = 0, n_skip = 0, n_calls = 0, n_gated = 0;
  for (u32 c4b = lo4; c4b <= hi4; c4b += 64u) {
    const u32 c4 = c4b + lane;
    const bool in = c4 <= hi4;
    const u32 n4 = in ? a.c6g[c4] : 0u;
    u32 t4 = xg + n4;
    t4 = min(t4, t4 - kGateMod);  // (T - c4^6) mod 124488
    bool live = in;
    if (a.use_tri) {
      const u32 y = t4 % 504u, z = t4 % 247u;
      live = in && ((s_tri[y >> 5] >> (y & 31u)) & 1u) && ((s_tri[16 + (z >> 5)] >> (z & 31u)) & 1u);
      n_c4 += (u64)__popcll(__ballot(in));
      n_skip += (u64)__popcll(__ballot(in && !live));
    }
    u32 l3 = 1, h3 = 0, x4 = 0;
    if (live) {
      const u128 c6 = p6_u128(c4);
      if (T > c6) {
        const u128 T4 = T - c6;
        const u32 h = min(iroot6_u128(T4), min(lim, c4));
        if (h >= 1) {
          const u32 l = min_core(T4, 3u);
          if (l <= h) {
            l3 = l;
            h3 = h;
            x4 = t4;
          }
        }
      }
    }
    u64 m = __ballot(l3 <= h3);
    while (m) {
      const int i = __builtin_ctzll(m);
      m &= m - 1;
      const u32 l = rdl(l3, i), h = rdl(h3, i), xi = rdl(x4, i);
      const u32 tag = (c4b + (u32)i) << 16;
      for (u32 cb = l; cb <= h; cb += 64u) {
        const u32 c = cb + lane;
        const bool inn = c <= h;
        const bool pass = inn && (!a.use_gate || gate_pass(s_gate, xi, a.c6g[c]));
        const u64 pm = __ballot(pass), im = __ballot(inn);
        n_calls += (u64)__popcll(pm);
        n_gated += (u64)__popcll(im & ~pm);
        tw_push(a, q, pass, ci, d, c | tag, lane);
        if (h - cb < 64u) break;  // cb + 64 could wrap past 2^32
      }
    }
    if (hi4 - c4b < 64u) break;
  }
  if (lane == 0) {
    TwinCounters* c = a.ctr;
    atomicAdd((unsigned long long*)&c->jobs, 1ULL);
    if (n_c4) atomicAdd((unsigned long long*)&c->c4_total, (unsigned long long)n_c4);
    if (n_skip) atomicAdd((unsigned long long*)&c->c4_tri_skipped, (unsigned long long)n_skip);
    if (n_calls) atomicAdd((unsigned long long*)&c->calls, (unsigned long long)n_calls);
    if (n_gated) atomicAdd((unsigned long long*)&c->gated, (unsigned long long)n_gated);
  }
}

__global__ __launch_bounds__(256) void k_twin_batched(TwinArgs a) {
  __shared__ u32 s_gate[kGateWords], s_tri[24];
  __shared__ u32 s_qc[kTwWaves][kTwRing], s_qd[kTwWaves][kTwRing], s_qx[kTwWaves][kTwRing];
  for (u32 i = threadIdx.x; i < kGateWords; i += blockDim.x) s_gate[i] = a.gate[i];
  if (threadIdx.x < 24) s_tri[threadIdx.x] = a.tri[threadIdx.x];
  __syncthreads();
  const u32 lane = threadIdx.x & 63u;
  const u32 w = threadIdx.x >> 6;
  TwRing q{s_qc[w], s_qd[w], s_qx[w], 0u, 0u};
  for (;;) {
    u32 item = 0;
    if (lane == 0) item = atomicAdd(a.item_ctr, 1u);
    item = __shfl(item, 0, 64);
    if (item >= a.n_items) break;
    twin_item(a, s_gate, s_tri, q, item, lane);
  }
  if (q.tail != q.head) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    tw_drain(a, q, q.tail - q.head, lane);
  }
}

}  // namespace

int launch_twin_batched(const TwinArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_twin_batched, dim3(grid), dim3(64 * kTwWaves), 0, s, a);
  return (int)hipGetLastError();
}

int twin_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_twin_batched, 64 * kTwWaves, 0) != hipSuccess) return 2;
  return nb > 0 ? nb : 1;
}

}  // namespace fc5
