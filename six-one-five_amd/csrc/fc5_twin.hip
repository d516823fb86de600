// fc5_twin.hip -- the (6,1,7) class-5 twin search on gfx950 (MI355X).
//
// k_twin_plain is the reference kernel shape for comparison: a block per
// (candidate, d, c4 tile of 2048), thread 0 computes the windows in R-space
// (double estimate + exact 128-bit fix-ups), the wavefronts stride c4, the
// lanes stride c3 and probe inline.
//
// k_twin_batched is a persistent kernel over a device work queue; one work
// item is one (candidate, d) cell of the reference's grid
// (k_find4_cls5, solve_617_v1.cu:563-657).  For its T it takes 64 values of
// c4 at a time, one per lane: triple-sum gate on T - c4^6 (bitmaps mod 504
// and 247 in LDS), the exact c3 window of T - c4^6 in 128 bits; then, one c4
// after another, the 64 lanes sweep c3 through the pair-sum gate (LDS bitmap) and compact the
// survivors into a per-wave LDS ring that is probed 64 keys at a time
// (key = T - c4^6 - c3^6 mod 2^64, 3 gathers each).  Every c4 of a window is
// visited once -- or, with a tile width, as often as the reference launch
// visits it: its c4 tiles start at lo4 + z * 2048 (z < zmax <= 32) and have
// no end (solve_617_v1.cu:607-608,961-979), so a c4 at offset k from lo4 is
// searched min(32, k / 2048 + 1) times; the counters and hits are multiplied
// accordingly.  Variant S (option strict_gate): the c3 rows run as a bit sieve
// (fc5_sieve.hpp: 32 consecutive c3 per lane, the pair-sum gate's five table
// words from LDS and the strict gate's ten from global memory, AND-ed), and only
// the strict survivors enter the ring.
#include "fc5_device.hpp"
#include "fc5_sieve.hpp"
#include "fc5_strict.hpp"
#include "fc5_twin.hpp"

namespace fc5 {

namespace {

constexpr int kTwWaves = 4;
__constant__ CSieveTab kTwCSieve = make_csieve();  // staged into LDS by the strict variant
constexpr u32 kTwRing = 128;

struct TwRing {
  u32 *qc, *qd, *qx;  // candidate index, d, c3 | c4 << 16
  u32 head, tail;
};

// visits of a c4 at offset k from the window start under the reference tiles
__device__ __forceinline__ u32 tile_visits(u32 k, u32 tile) { return tile ? min(32u, k / tile + 1u) : 1u; }

__device__ __forceinline__ void tw_drain(const TwinArgs& a, TwRing& q, u32 n, u32 lane) {
  if (lane >= n) return;
  const u32 slot = (q.head + lane) & (kTwRing - 1);
  const u32 ci = q.qc[slot], d = q.qd[slot], x = q.qx[slot];
  const u32 c3 = x & 0xffffu, c4 = x >> 16;
  const DevTwinCand* C = a.cands + ci;
  const u64 d3 = (u64)d * d * d;
  const u128 dq = ((u128)d3 * d3) >> 6;
  const u64 a2 = (u64)c4 * c4, b2 = (u64)c3 * c3;
  const u64 fp = C->Q_lo - (u64)dq - a2 * a2 * a2 - b2 * b2 * b2;
  if (d_query(a.main, fp)) {
    u32 visits = 1;
    if (a.tile) {  // rare: recompute the window start of this (candidate, d)
      const u128 T = (((u128)C->Q_hi << 64) | C->Q_lo) - dq;
      visits = tile_visits(c4 - min_core(T, a.c4_div), a.tile);
    }
    atomicAdd((unsigned long long*)&a.ctr->maybe, (unsigned long long)visits);
    const u64 s = atomicAdd((unsigned long long*)a.nhits, (unsigned long long)visits);
    for (u32 v = 0; v < visits; ++v)
      if (s + v < a.hit_cap) {
        TwinHit h;
        h.cand = C->uid;
        h.d = d;
        h.c4 = c4;
        h.c3 = c3;
        h.fp = fp;
        a.hits[s + v] = h;
      }
  }
}

__device__ __forceinline__ void tw_push(const TwinArgs& a, TwRing& q, bool pass, u32 ci, u32 d, u32 x, u32 lane) {
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

// variant S: the sieve's per-lane counts (calls, gated, queries), reduced at the end of the launch
struct TwLaneCounts {
  u64 calls, gated, queries;
};

template <bool S>
__device__ __forceinline__ void twin_item(const TwinArgs& a, const u32* s_gate, const u32* s_tri,
                                          const u32* s_csieve, TwRing& q, TwLaneCounts& lc, u32 item, u32 lane) {
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
    const u32 n4 = in ? neg_p6_gate(c4) : 0u;  // c4 <= lim < 2^16
    u32 t4 = xg + n4;
    t4 = min(t4, t4 - kGateMod);  // (T - c4^6) mod 124488
    bool live = in;
    const u32 visits = tile_visits(c4 - lo4, a.tile);  // 1 unless the reference tiles are asked for
    if (a.use_tri) {
      const u32 y = t4 % 504u, z = t4 % 247u;
      live = in && ((s_tri[y >> 5] >> (y & 31u)) & 1u) && ((s_tri[16 + (z >> 5)] >> (z & 31u)) & 1u);
      if (a.tile) {
        n_c4 += wave_sum(in ? visits : 0u);
        n_skip += wave_sum(in && !live ? visits : 0u);
      } else {
        n_c4 += (u64)__popcll(__ballot(in));
        n_skip += (u64)__popcll(__ballot(in && !live));
      }
    }
    u32 l3 = 1, h3 = 0, x4 = 0;
    StrictRow s4{0u, 0u};
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
            if (S) s4 = strict_row((u64)(T4 >> 64), (u64)T4);
          }
        }
      }
    }
    u64 m = __ballot(l3 <= h3);
    while (m) {
      const int i = __builtin_ctzll(m);
      m &= m - 1;
      const u32 l = rdl(l3, i), h = rdl(h3, i), xi = rdl(x4, i), vi = rdl(visits, i);
      const u32 tag = (c4b + (u32)i) << 16;
      if (S) {
        u32 ra, rb, rc;
        csieve_row_bases(xi, ra, rb, rc);
        const StrictSieveRow sb = strict_csieve_bases(StrictRow{rdl(s4.w0, i), rdl(s4.w1, i)});
        const u32 nw = (h - l) / 32u + 1u;
        for (u32 w0 = 0; w0 < nw; w0 += 64u) {
          const u32 w = w0 + lane;
          const u32 c0 = l + 32u * w;  // below 2^17: l, h <= lim < 2^16
          u32 P = 0, cut = 0;
          if (w < nw) {
            cut = csieve_cut(c0, h);
            P = a.use_gate ? cut & csieve_word(s_csieve, ra, rb, rc, c0) : cut;
          }
          lc.calls += (u64)__popc(P) * vi;
          lc.gated += (u64)(__popc(cut) - __popc(P)) * vi;
          if (P) P &= strict_csieve_word(a.scsieve, sb, c0);
          lc.queries += (u64)__popc(P) * vi;
          while (__ballot(P != 0u)) {
            const bool has = P != 0u;
            const u32 c = c0 + (has ? (u32)__builtin_ctz(P) : 0u);
            tw_push(a, q, has, ci, d, c | tag, lane);
            P &= P - 1u;
          }
        }
        continue;
      }
      for (u32 cb = l; cb <= h; cb += 64u) {
        const u32 c = cb + lane;
        const bool inn = c <= h;
        const bool pass = inn && (!a.use_gate || gate_pass(s_gate, xi, neg_p6_gate(c)));
        const u64 pm = __ballot(pass), im = __ballot(inn);
        n_calls += (u64)__popcll(pm) * vi;
        n_gated += (u64)__popcll(im & ~pm) * vi;
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

template <bool S>
__global__ __launch_bounds__(256) void k_twin_batched(TwinArgs a) {
  __shared__ u32 s_gate[kGateWords], s_tri[24], s_csieve[S ? kCSieveWords : 1];
  __shared__ u32 s_qc[kTwWaves][kTwRing], s_qd[kTwWaves][kTwRing], s_qx[kTwWaves][kTwRing];
  for (u32 i = threadIdx.x; i < kGateWords; i += blockDim.x) s_gate[i] = a.gate[i];
  if (threadIdx.x < 24) s_tri[threadIdx.x] = a.tri[threadIdx.x];
  if (S)
    for (u32 i = threadIdx.x; i < kCSieveWords; i += blockDim.x) s_csieve[i] = kTwCSieve.w[i];
  __syncthreads();
  const u32 lane = threadIdx.x & 63u;
  const u32 w = threadIdx.x >> 6;
  TwRing q{s_qc[w], s_qd[w], s_qx[w], 0u, 0u};
  TwLaneCounts lc{0, 0, 0};
  for (;;) {
    u32 item = 0;
    if (lane == 0) item = atomicAdd(a.item_ctr, 1u);
    item = __shfl(item, 0, 64);
    if (item >= a.n_items) break;
    twin_item<S>(a, s_gate, s_tri, s_csieve, q, lc, item, lane);
  }
  if (q.tail != q.head) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    tw_drain(a, q, q.tail - q.head, lane);
  }
  if (S) {
    const u64 calls = wave_sum(lc.calls), gated = wave_sum(lc.gated), queries = wave_sum(lc.queries);
    if (lane == 0) {
      if (calls) atomicAdd((unsigned long long*)&a.ctr->calls, (unsigned long long)calls);
      if (gated) atomicAdd((unsigned long long*)&a.ctr->gated, (unsigned long long)gated);
      if (queries) atomicAdd((unsigned long long*)&a.ctr->queries, (unsigned long long)queries);
    }
  }
}

// ---------------------------------------------------------------------------
// plain form: the reference kernel shape (k_find4_cls5, solve_617_v1.cu:563-657)
// ---------------------------------------------------------------------------
constexpr u64 kM42 = 5489031744ULL;

// s * c^6 > R (s * c^6 past 2^128 counts as greater)
__device__ __forceinline__ bool scaled_gt(u128 R, u32 c, u64 s) {
  u128 v;
  if (__builtin_mul_overflow(p6_u128(c), (u128)s, &v)) return true;
  return v > R;
}

// largest r with s * r^6 <= R, from a double estimate and exact fix-ups
__device__ u32 root_fix(u128 R, u64 s) {
  const double x = ((double)(u64)(R >> 64) * 18446744073709551616.0 + (double)(u64)R) / (double)s;
  u32 r = (u32)pow(x, 1.0 / 6.0) + 2;
  while (r > 0 && scaled_gt(R, r, s)) --r;
  while (!scaled_gt(R, r + 1, s)) ++r;
  return r;
}

// least c >= 1 with R <= s * c^6
__device__ u32 min_fix(u128 R, u32 start, u64 s) {
  u32 c = start < 1 ? 1 : start;
  while (!scaled_gt(R, c, s) && (u128)s * p6_u128(c) != R) ++c;
  while (c > 1 && (scaled_gt(R, c - 1, s) || (u128)s * p6_u128(c - 1) == R)) --c;
  return c;
}

// (hi:lo) mod m with c = 2^64 mod m, m < 2^42
__device__ __forceinline__ u64 mod128(u128 R, u64 m, u64 c) {
  const u64 hi = (u64)(R >> 64), lo = (u64)R;
  const u64 r = hi % m;
  const u128 rc = (u128)r * c;
  u64 t = (u64)((u128)(u64)(rc >> 64) * c % m);
  t += (u64)((u64)rc % m);
  if (t >= m) t -= m;
  t += lo % m;
  if (t >= m) t -= m;
  return t;
}

__global__ __launch_bounds__(256) void k_twin_plain(PlainTwinArgs a) {
  const u32 ci = blockIdx.x;
  if (ci >= a.n_cands) return;
  const PlainTwinCand& C = a.cands[ci];
  const u32 kmax = C.fmax1 / 64u + 1u;
  const u32 t = blockIdx.y;
  if (t >= C.nres * kmax) return;
  const u32 d = C.res[t % C.nres] + (t / C.nres) * 64u;
  if (d == 0 || d > C.fmax1) return;
  __shared__ u64 s_rlo, s_rhi;
  __shared__ u32 s_lo4, s_hi4, s_q504, s_q247;
  __shared__ int s_active;
  __shared__ u32 s_tri[24];
  __shared__ u16 s_p504[504], s_p247[247];
  if (threadIdx.x == 0) {
    s_active = 0;
    const u128 base = ((u128)C.b_hi << 64) | C.b_lo;
    const u128 g6 = p6_u128(21u * d);
    if (base > g6) {
      const u128 R = base - g6;
      u32 hi4 = root_fix(R, kM42);
      if (hi4 > C.lim) hi4 = C.lim;
      if (hi4 >= 1) {
        const u32 lo4 = min_fix(R, (u32)(hi4 * 0.8), (u64)a.c4_div * kM42);
        if (lo4 <= hi4) {
          s_rlo = (u64)R;
          s_rhi = (u64)(R >> 64);
          s_lo4 = lo4;
          s_hi4 = hi4;
          s_q504 = (u32)(mod128(R, 504ULL * kM42, a.c_m504) / kM42 % 504ULL);
          s_q247 = (u32)(mod128(R, 247ULL, a.c_m247) * a.inv42_247 % 247ULL);
          s_active = 1;
          // a block per (cand, d, tile): count the cell once, in tile 0
          if (blockIdx.z == 0) atomicAdd((unsigned long long*)&a.ctr->jobs, 1ULL);
        }
      }
    }
  }
  __syncthreads();
  if (!s_active) return;
  for (u32 i = threadIdx.x; i < 504; i += blockDim.x) s_p504[i] = a.p6[i];
  for (u32 i = threadIdx.x; i < 247; i += blockDim.x) s_p247[i] = a.p6[504 + i];
  if (threadIdx.x < 24) s_tri[threadIdx.x] = a.tri[threadIdx.x];
  __syncthreads();
  const u128 R = ((u128)s_rhi << 64) | s_rlo;
  const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  // tiles of 2048 with an end (every c4 once), or the reference's unbounded tiles
  const u32 t0 = s_lo4 + blockIdx.z * (a.tile ? a.tile : 2048u);
  const u32 t1 = a.tile ? s_hi4 : min(s_hi4, t0 + 2047u);
  u64 n_c4 = 0, n_skip = 0, n_calls = 0, n_gated = 0;
  for (u32 c4 = t0 + wave; c4 <= t1; c4 += waves) {
    u32 t4504 = s_q504 + 504u - s_p504[c4 % 504u];
    if (t4504 >= 504u) t4504 -= 504u;
    u32 t4247 = s_q247 + 247u - s_p247[c4 % 247u];
    if (t4247 >= 247u) t4247 -= 247u;
    if (a.use_tri) {
      ++n_c4;
      if (!((s_tri[t4504 >> 5] >> (t4504 & 31u)) & 1u) || !((s_tri[16 + (t4247 >> 5)] >> (t4247 & 31u)) & 1u)) {
        ++n_skip;
        continue;
      }
    }
    u128 m4;
    if (__builtin_mul_overflow(p6_u128(c4), (u128)kM42, &m4) || m4 >= R) continue;
    const u128 R4 = R - m4;
    u32 hi3 = root_fix(R4, kM42);
    hi3 = min(hi3, min(C.lim, c4));
    if (hi3 < 1) continue;
    const u32 lo3 = min_fix(R4, (u32)(hi3 * 0.83), 3ULL * kM42);
    const u64 tfp = (((u64)R4 >> 6) | ((u64)(R4 >> 64) << 58)) * a.inv216;  // T4 mod 2^64
    u32 r504 = (lo3 + lane) % 504u, r247 = (lo3 + lane) % 247u;
    for (u32 c3 = lo3 + lane; c3 <= hi3; c3 += 64u) {
      bool pass = true;
      if (a.use_gate) {
        u32 x = t4504 + 504u - s_p504[r504];
        if (x >= 504u) x -= 504u;
        u32 y = t4247 + 247u - s_p247[r247];
        if (y >= 247u) y -= 247u;
        pass = ((a.gate.g504[x >> 6] >> (x & 63u)) & 1ULL) && ((a.gate.g247[y >> 6] >> (y & 63u)) & 1ULL);
        r504 += 64u;
        if (r504 >= 504u) r504 -= 504u;
        r247 += 64u;
        if (r247 >= 247u) r247 -= 247u;
      }
      if (!pass) {
        ++n_gated;
        continue;
      }
      ++n_calls;
      const u64 c2 = (u64)c3 * c3;
      const u64 fp = tfp - c2 * c2 * c2;
      if (d_query(a.main, fp)) {
        atomicAdd((unsigned long long*)&a.ctr->maybe, 1ULL);
        const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
        if (s < a.hit_cap) {
          TwinHit h;
          h.cand = C.uid;
          h.d = d;
          h.c4 = c4;
          h.c3 = c3;
          h.fp = fp;
          a.hits[s] = h;
        }
      }
      if (hi3 - c3 < 64u) break;
    }
  }
  n_calls = wave_sum(n_calls);
  n_gated = wave_sum(n_gated);
  if (lane == 0) {
    TwinCounters* c = a.ctr;
    if (n_c4) atomicAdd((unsigned long long*)&c->c4_total, (unsigned long long)n_c4);
    if (n_skip) atomicAdd((unsigned long long*)&c->c4_tri_skipped, (unsigned long long)n_skip);
    if (n_calls) atomicAdd((unsigned long long*)&c->calls, (unsigned long long)n_calls);
    if (n_gated) atomicAdd((unsigned long long*)&c->gated, (unsigned long long)n_gated);
  }
}

}  // namespace

int launch_twin_plain(const PlainTwinArgs& a, u32 ymax, u32 zmax, hipStream_t s) {
  hipLaunchKernelGGL(k_twin_plain, dim3(a.n_cands, ymax, zmax), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

int launch_twin_batched(const TwinArgs& a, int grid, hipStream_t s) {
  if (a.strict) hipLaunchKernelGGL(k_twin_batched<true>, dim3(grid), dim3(64 * kTwWaves), 0, s, a);
  else hipLaunchKernelGGL(k_twin_batched<false>, dim3(grid), dim3(64 * kTwWaves), 0, s, a);
  return (int)hipGetLastError();
}

int twin_blocks_per_cu(bool strict) {
  int nb = 0;
  const hipError_t e = strict
                           ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_twin_batched<true>, 64 * kTwWaves, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_twin_batched<false>, 64 * kTwWaves, 0);
  if (e != hipSuccess) return 2;
  return nb > 0 ? nb : 1;
}

}  // namespace fc5
