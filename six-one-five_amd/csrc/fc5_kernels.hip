// fc5_kernels.hip -- class-5 free-peel kernels for gfx950 (MI355X).
//
// k_cls5_batched  persistent kernel fed by a device work queue (one item =
//                 one unit x 8 e-rows x all d).  Per point it tests the
//                 sixth-power residue gate in T-space with 32-bit VALU work
//                 and an LDS bitmap; survivors are compacted per wavefront
//                 (ballot + mbcnt into an LDS ring) and probed 64 at a time:
//                 fingerprint, 3 mixes, 3 exact remainders, 3 HBM gathers.
// k_cls5_plain    the reference kernel shape (fourcore_cls5_gpu_v4.cu:353-410):
//                 block per (unit, e), threads over d, per-point 192-bit
//                 sixth power, subtract-with-borrow and modular gate.
// k_expand_cells  widens the r-bit packed table to one u64 per cell.
// k_query         membership queries for a key list (tests, FPR smoke).
#include "fc5_kernels.hpp"

namespace fc5 {

typedef unsigned __int128 u128;

constexpr u64 kM42 = 5489031744ULL;
constexpr u32 kGateMod = 124488;
constexpr u32 kGateWords = (kGateMod + 31) / 32;
constexpr int kWaves = 4;   // wavefronts per workgroup
constexpr u32 kRing = 128;  // per-wave survivor ring (entries)

__device__ __forceinline__ u64 d_mix64(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// xor filter query on expanded cells (xor_filter.hpp:53-64,472-483)
__device__ __forceinline__ bool d_query(const XorDev& x, u64 key) {
  if (!x.present) return false;
  const u64 h = d_mix64(key + x.seed);
  const u64 h1 = d_mix64(h ^ 0x9E3779B97F4A7C15ULL);
  const u64 h2 = d_mix64(h ^ 0xBF58476D1CE4E5B9ULL);
  const u64 i0 = mod_fast(h, x.md);
  const u64 i1 = mod_fast(h1, x.md) + x.block;
  const u64 i2 = mod_fast(h2, x.md) + 2 * x.block;
  const u64 got = load_cell_policy(x.cells + i0, x.policy) ^ load_cell_policy(x.cells + i1, x.policy) ^
                  load_cell_policy(x.cells + i2, x.policy);
  return got == (h & x.mask);
}

__device__ __forceinline__ u64 wave_sum(u64 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ u32 rdl(u32 v, int lane) {
  return (u32)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ u32 lanes_below(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// floor(x^6 / 64) for x < 2^21
__device__ __forceinline__ u128 dq_of(u32 x) {
  const u64 x3 = (u64)x * x * x;
  return ((u128)x3 * x3) >> 6;
}

// ---------------------------------------------------------------------------
// batched persistent kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void drain(const BatchedArgs& a, const u32* qu, const u32* qd,
                                      const u32* qe, u32 head, u32 n, u32 lane) {
  if (lane < n) {
    const u32 slot = (head + lane) & (kRing - 1);
    const u32 ui = qu[slot], d = qd[slot], e = qe[slot];
    const DevUnit5* U = a.units + ui;
    const u64 e2 = (u64)e * e;
    const u64 eq = (e2 * e2 * e2 - (u64)U->te) * a.inv729;  // floor(e^6/729) mod 2^64
    const u64 d3 = (u64)d * d * d;
    const u64 dq = (u64)(((u128)d3 * d3) >> 6);             // floor(d^6/64) mod 2^64
    const u64 fp = U->Q_lo - eq - dq;                        // T mod 2^64
    u32 flags = d_query(a.main, fp) ? 1u : 0u;
    if (a.side.present && d_query(a.side, fp)) flags |= 2u;
    if (flags) {
      Counters* c = a.ctr + U->band;
      if (flags & 1u) atomicAdd((unsigned long long*)&c->maybe, 1ULL);
      if (flags & 2u) atomicAdd((unsigned long long*)&c->side_maybe, 1ULL);
      const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
      if (s < a.hit_cap) {
        HitRec h;
        h.unit = U->uid;
        h.d = d;
        h.e = e;
        h.flags = flags;
        h.fp = fp;
        a.hits[s] = h;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_cls5_batched(BatchedArgs a) {
  __shared__ u32 s_gate[kGateWords];
  __shared__ u32 s_qu[kWaves][kRing], s_qd[kWaves][kRing], s_qe[kWaves][kRing];
  for (u32 i = threadIdx.x; i < kGateWords; i += blockDim.x) s_gate[i] = a.gate[i];
  __syncthreads();

  const u32 lane = threadIdx.x & 63u;
  const u32 w = threadIdx.x >> 6;
  u32* qu = s_qu[w];
  u32* qd = s_qd[w];
  u32* qe = s_qe[w];
  u32 head = 0, tail = 0;
  const u32 K = a.ndg_K;

  for (;;) {
    u32 item = 0;
    if (lane == 0) item = atomicAdd(a.item_ctr, 1u);
    item = __shfl(item, 0, 64);
    if (item >= a.n_items) break;

    // unit of this item: last unit with item_base <= item
    u32 lo = 0, hi = a.n_units;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (a.units[mid].item_base <= item) lo = mid;
      else hi = mid;
    }
    const u32 ui = lo;
    const DevUnit5 U = a.units[ui];
    const u32 row0 = (item - U.item_base) * (u32)kRowsPerItem;
    const u32 r0 = U.res2[0], r1 = U.res2[1], r2 = U.res2[2], r3 = U.res2[3];
    const u32 minr = min(min(r0, r1), min(r2, r3));

    // ---- per-row set-up on lanes 0..7: e, T-space row value, positive cut
    u32 e = 0, dlim = 0, xg = 0;
    u32 c_calls = 0, c_neg = 0, c_zero = 0, c_pos = 0;
    if (lane < (u32)kRowsPerItem) {
      const u32 r = row0 + lane;
      if (r < U.rows) {
        e = (u32)a.units[ui].res1[r % 6u] + (r / 6u) * 729u;
        if (e >= 1 && e <= U.fmax1) {
          auto cnt = [](u32 rj, u32 lim) -> u32 { return rj <= lim ? (lim - rj) / 64u + 1u : 0u; };
          const u32 cr = cnt(r0, U.fmax2) + cnt(r1, U.fmax2) + cnt(r2, U.fmax2) + cnt(r3, U.fmax2);
          const u64 e3 = (u64)e * e * e;
          const u128 eq = (((u128)e3 * e3) - U.te) / 729u;
          const u128 Q = ((u128)U.Q_hi << 64) | U.Q_lo;
          u32 pos = 0, zero = 0;
          if (Q > eq) {
            const u128 X = Q - eq;  // T = X - floor(d^6/64)
            const double xd = ((double)(u64)(X >> 64) * 18446744073709551616.0 + (double)(u64)X) * 64.0;
            u32 dc = (u32)cbrt(sqrt(xd));
            while (dq_of(dc + 1) < X) ++dc;
            while (dc > 0 && dq_of(dc) >= X) --dc;
            dlim = min(dc, U.fmax2);
            pos = cnt(r0, dlim) + cnt(r1, dlim) + cnt(r2, dlim) + cnt(r3, dlim);
            const u64 xh = (u64)(X >> 64), xl = (u64)X;
            xg = (u32)(((xh % kGateMod) * (u64)(18446744073709551615ULL % kGateMod + 1) +
                        xl % kGateMod) % kGateMod);
            const u32 z = dc + 1;  // the only d that can give T == 0
            if (z <= U.fmax2 && dq_of(z) == X) {
              const u32 zr = z & 63u;
              zero = (zr == r0 || zr == r1 || zr == r2 || zr == r3) ? 1u : 0u;
            }
          } else if (Q == eq) {
            zero = (U.fmax2 >= 1 && (r0 == 1 || r1 == 1 || r2 == 1 || r3 == 1)) ? 1u : 0u;
          }
          c_calls = cr;
          c_pos = pos;
          c_zero = zero;
          c_neg = cr - pos - zero;
        }
      }
    }
    u32 maxdl = 0;
    u64 s_calls = 0, s_neg = 0, s_zero = 0, s_pos = 0;
#pragma unroll
    for (int i = 0; i < kRowsPerItem; ++i) {
      maxdl = max(maxdl, rdl(dlim, i));
      s_calls += rdl(c_calls, i);
      s_neg += rdl(c_neg, i);
      s_zero += rdl(c_zero, i);
      s_pos += rdl(c_pos, i);
    }

    // ---- gate over the positive part of the rows; lanes run over k
    u32 probes_item = 0;
    for (u32 kb = 0; minr + 64u * kb <= maxdl; kb += 64u) {
      const u32 k = kb + lane;
      const u32 d0 = r0 + 64u * k, d1 = r1 + 64u * k, d2 = r2 + 64u * k, d3 = r3 + 64u * k;
      const u32 n0 = d0 <= U.fmax2 ? a.ndg[r0 * K + k] : 0u;
      const u32 n1 = d1 <= U.fmax2 ? a.ndg[r1 * K + k] : 0u;
      const u32 n2 = d2 <= U.fmax2 ? a.ndg[r2 * K + k] : 0u;
      const u32 n3 = d3 <= U.fmax2 ? a.ndg[r3 * K + k] : 0u;
      const u32 dmin_blk = minr + 64u * kb;
      for (int i = 0; i < kRowsPerItem; ++i) {
        const u32 dl = rdl(dlim, i);
        if (dl < dmin_blk) continue;
        const u32 xgi = rdl(xg, i);
        const u32 ei = rdl(e, i);
        const u32 ds[4] = {d0, d1, d2, d3};
        const u32 ns[4] = {n0, n1, n2, n3};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          u32 t = xgi + ns[j];
          t = min(t, t - kGateMod);
          const bool pass = ((s_gate[t >> 5] >> (t & 31u)) & 1u) && ds[j] <= dl;
          const u64 m = __ballot(pass);
          if (m) {
            if (pass) {
              const u32 slot = (tail + lanes_below(m)) & (kRing - 1);
              qu[slot] = ui;
              qd[slot] = ds[j];
              qe[slot] = ei;
            }
            const u32 cnt = (u32)__popcll(m);
            tail += cnt;
            probes_item += cnt;
            if (tail - head >= 64u) {
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
              drain(a, qu, qd, qe, head, 64u, lane);
              head += 64u;
            }
          }
        }
      }
    }
    if (lane == 0) {
      Counters* c = a.ctr + U.band;
      atomicAdd((unsigned long long*)&c->calls, (unsigned long long)s_calls);
      atomicAdd((unsigned long long*)&c->neg, (unsigned long long)s_neg);
      atomicAdd((unsigned long long*)&c->zero, (unsigned long long)s_zero);
      atomicAdd((unsigned long long*)&c->pos, (unsigned long long)s_pos);
      atomicAdd((unsigned long long*)&c->probes, (unsigned long long)probes_item);
    }
  }
  if (tail != head) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    drain(a, qu, qd, qe, head, tail - head, lane);
  }
}

// ---------------------------------------------------------------------------
// plain kernel: the reference's per-point 192-bit arithmetic and launch shape
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pow6_192(u32 x, u64& o2, u64& o1, u64& o0) {
  const u64 x2 = (u64)x * x;
  const u128 x4 = (u128)x2 * x2;
  const u128 lo = (u128)(u64)x4 * x2;
  const u128 hi = (u128)(u64)(x4 >> 64) * x2;
  o0 = (u64)lo;
  const u128 mid = (lo >> 64) + (u128)(u64)hi;
  o1 = (u64)mid;
  o2 = (u64)(hi >> 64) + (u64)(mid >> 64);
}

__device__ __forceinline__ bool sub192(u64& a2, u64& a1, u64& a0, u64 b2, u64 b1, u64 b0) {
  const u128 s0 = (u128)a0 - b0;
  const u64 br0 = (u64)(s0 >> 64) & 1u;
  const u128 s1 = (u128)a1 - b1 - br0;
  const u64 br1 = (u64)(s1 >> 64) & 1u;
  const u128 s2 = (u128)a2 - b2 - br1;
  a0 = (u64)s0;
  a1 = (u64)s1;
  a2 = (u64)s2;
  return ((u64)(s2 >> 64) & 1u) != 0;
}

__device__ __forceinline__ u64 mod192(u64 a2, u64 a1, u64 a0, u64 m, u64 two64) {
  u64 r = a2 % m;
  r = (u64)((u128)r * two64 % m);
  r = (u64)((u128)r * two64 % m);
  r = (u64)(((u128)r + (u128)(a1 % m) * two64) % m);
  r += a0 % m;
  if (r >= m) r -= m;
  return r;
}

__global__ __launch_bounds__(256) void k_cls5_plain(PlainArgs a) {
  const int ui = blockIdx.x;
  if (ui >= a.n_units) return;
  const PlainUnit& U = a.units[ui];
  const u32 kmax1 = U.fmax1 / U.mod1 + 1;
  const u32 t1 = blockIdx.y;
  if (t1 >= U.nres1 * kmax1) return;
  const u32 e = U.res[t1 % U.nres1] + (t1 / U.nres1) * U.mod1;
  if (e == 0 || e > U.fmax1) return;
  __shared__ u64 s_e[3];
  if (threadIdx.x == 0) pow6_192(14u * e, s_e[2], s_e[1], s_e[0]);
  __syncthreads();
  const u32 n2 = U.nres2, total = n2 * (U.fmax2 / U.mod2 + 1);
  u64 calls = 0, neg = 0, zero = 0, gated = 0, probes = 0;
  for (u32 t2 = threadIdx.x; t2 < total; t2 += blockDim.x) {
    const u32 d = U.res[U.nres1 + t2 % n2] + (t2 / n2) * U.mod2;
    if (d == 0 || d > U.fmax2) continue;
    ++calls;
    u64 x0 = U.b0, x1 = U.b1, x2 = U.b2;
    if (sub192(x2, x1, x0, s_e[2], s_e[1], s_e[0])) { ++neg; continue; }
    u64 p2, p1, p0;
    pow6_192(21u * d, p2, p1, p0);
    if (sub192(x2, x1, x0, p2, p1, p0)) { ++neg; continue; }
    if ((x0 | x1 | x2) == 0) { ++zero; continue; }
    if (a.use_gate) {
      const u64 rm = mod192(x2, x1, x0, 504ULL * kM42, a.c_m504);
      const u32 t504 = (u32)((rm / kM42) % 504ULL);
      const u64 r247 = mod192(x2, x1, x0, 247ULL, a.c_m247);
      const u32 t247 = (u32)(r247 * a.inv42_247 % 247ULL);
      const bool ok = ((a.gate.g504[t504 >> 6] >> (t504 & 63)) & 1ULL) &&
                      ((a.gate.g247[t247 >> 6] >> (t247 & 63)) & 1ULL);
      if (!ok) { ++gated; continue; }
    }
    ++probes;
    const u64 fp = ((x0 >> 6) | (x1 << 58)) * a.inv216;
    u32 flags = d_query(a.main, fp) ? 1u : 0u;
    if (a.side.present && d_query(a.side, fp)) flags |= 2u;
    if (flags) {
      if (flags & 1u) atomicAdd((unsigned long long*)&a.ctr->maybe, 1ULL);
      if (flags & 2u) atomicAdd((unsigned long long*)&a.ctr->side_maybe, 1ULL);
      const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
      if (s < a.hit_cap) {
        HitRec h;
        h.unit = U.uid;
        h.d = d;
        h.e = e;
        h.flags = flags;
        h.fp = fp;
        a.hits[s] = h;
      }
    }
  }
  calls = wave_sum(calls);
  neg = wave_sum(neg);
  zero = wave_sum(zero);
  gated = wave_sum(gated);
  probes = wave_sum(probes);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd((unsigned long long*)&a.ctr->calls, (unsigned long long)calls);
    atomicAdd((unsigned long long*)&a.ctr->neg, (unsigned long long)neg);
    atomicAdd((unsigned long long*)&a.ctr->zero, (unsigned long long)zero);
    atomicAdd((unsigned long long*)&a.ctr->gated, (unsigned long long)gated);
    atomicAdd((unsigned long long*)&a.ctr->probes, (unsigned long long)probes);
  }
}

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
__global__ void k_expand_cells(const u32* __restrict__ pw, u64 m, u32 r, u64* __restrict__ cells) {
  const u64 mask = r >= 64 ? ~0ULL : ((1ULL << r) - 1ULL);
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (u64)gridDim.x * blockDim.x) {
    const u64 bit = i * r;
    const u64 wi = bit >> 5;
    const u32 sh = (u32)(bit & 31u);
    u64 v = ((u64)pw[wi] | ((u64)pw[wi + 1] << 32)) >> sh;
    if (sh + r > 64u) v |= (u64)pw[wi + 2] << (64u - sh);
    cells[i] = v & mask;
  }
}

__global__ void k_query(XorDev x, const u64* __restrict__ keys, u64 n, u8* __restrict__ out) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    out[i] = d_query(x, keys[i]) ? 1 : 0;
}

__global__ void k_fill_u64(u64* p, u64 n, u64 v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    p[i] = v;
}

static int grid_for(u64 n, int threads) {
  const u64 g = (n + (u64)threads - 1) / (u64)threads;
  return (int)(g < 65536 ? (g ? g : 1) : 65536);
}

int launch_cls5_batched(const BatchedArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_cls5_batched, dim3(grid), dim3(64 * kWaves), 0, s, a);
  return (int)hipGetLastError();
}

int launch_cls5_plain(const PlainArgs& a, int n_units, int ymax, hipStream_t s) {
  hipLaunchKernelGGL(k_cls5_plain, dim3(n_units, ymax), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

int launch_expand_cells(const u32* pw, u64 m, u32 r, u64* cells, hipStream_t s) {
  hipLaunchKernelGGL(k_expand_cells, dim3(grid_for(m, 256)), dim3(256), 0, s, pw, m, r, cells);
  return (int)hipGetLastError();
}

int launch_query(const XorDev& x, const u64* keys, u64 n, u8* out, hipStream_t s) {
  hipLaunchKernelGGL(k_query, dim3(grid_for(n, 256)), dim3(256), 0, s, x, keys, n, out);
  return (int)hipGetLastError();
}

int launch_fill_u64(u64* p, u64 n, u64 v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
  return (int)hipGetLastError();
}

int batched_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cls5_batched, 64 * kWaves, 0) != hipSuccess)
    return 4;
  return nb > 0 ? nb : 1;
}

}  // namespace fc5
