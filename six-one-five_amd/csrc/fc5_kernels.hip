// fc5_kernels.hip -- class-5 free-peel kernels for gfx950 (MI355X).
//
// k_cls5_batched  persistent kernel fed by a device work queue (one item =
//                 one unit x E e-rows x all d, or E d-rows of a class 2-4
//                 unit; the launch's items are segments of unit arenas, so
//                 one launch can serve several bands).  Variant V (option
//                 sweep = 1, the default) tests the sixth-power residue gate in
//                 T-space as a bit sieve, 32 grid points per lane and LDS table
//                 word (fc5_sieve.hpp); variant !V tests one point per lane
//                 against an LDS bitmap.  Survivors are compacted per wavefront
//                 (ballot + mbcnt into an LDS ring) and probed 64 at a time:
//                 fingerprint, 3 mixes, 3 exact remainders, 3 HBM gathers.
//                 Variant S (option strict_gate): V applies the strict gate as
//                 ten more sieve words and pushes its survivors straight to
//                 the probe ring; !V tests each batch of 64 gate survivors
//                 (fc5_strict.hpp) and compacts the passing ones into it.
// k_cls5_plain    the reference kernel shape (fourcore_cls5_gpu_v4.cu:353-410):
//                 block per (unit, e), threads over d, per-point 192-bit
//                 sixth power, subtract-with-borrow and modular gate.
// k_expand_cells  widens the r-bit packed table to one u64 per cell.
// k_query         membership queries for a key list (tests, FPR smoke).
#include <vector>

#include "fc5_device.hpp"
#include "fc5_sieve.hpp"
#include "fc5_strict.hpp"

namespace fc5 {

constexpr u64 kM42 = 5489031744ULL;
constexpr int kWaves = 4;   // wavefronts per workgroup
constexpr u32 kRing = 128;  // per-wave survivor ring (entries)

__constant__ SieveTab kSieveDev = make_sieve();     // staged into LDS by every workgroup
__constant__ CSieveTab kCSieveDev = make_csieve();  // the same, consecutive c (class 2-4 rows)
__device__ u32 g_strict_sieve[kStrictWords];        // fill_strict_sieve, uploaded per device (upload_sieve_tables)
__device__ u32 g_strict_csieve[kStrictCWords];      // fill_strict_csieve, likewise

// floor(x^6 / 64) for x < 2^21
__device__ __forceinline__ u128 dq_of(u32 x) {
  const u64 x3 = (u64)x * x * x;
  return ((u128)x3 * x3) >> 6;
}

// ---------------------------------------------------------------------------
// batched persistent kernel
// ---------------------------------------------------------------------------
constexpr u32 kFind3Bit = 0x80000000u;  // ring entry unit field: find3 row

// A wavefront's counter sums not yet added to its band's counters: one band at
// a time, flushed when the band changes and at the end of the launch, so the
// counters take a few atomics per wavefront instead of several per work item.
struct WaveAcc {
  u64 v[8];  // calls, neg, zero, pos, probes, jobs, failT, queries
  u32 band, live;
  u32 chunk, cfill;  // probe mode 1: the wavefront's emit chunk (~0u: none yet) and its fill
};

// lane 0: add the sums to the band's counters and clear them
__device__ __forceinline__ void acc_flush(const BatchedArgs& a, WaveAcc* A) {
  if (!A->live) return;
  Counters* c = a.ctr + A->band;
  u64* dst[8] = {&c->calls, &c->neg, &c->zero, &c->pos, &c->probes, &c->jobs, &c->failT, &c->queries};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (A->v[i]) atomicAdd((unsigned long long*)dst[i], (unsigned long long)A->v[i]);
    A->v[i] = 0;
  }
  A->live = 0;
}

// lane 0: make `band` the accumulated one
__device__ __forceinline__ void acc_band(const BatchedArgs& a, WaveAcc* A, u32 band) {
  if (A->live && A->band != band) acc_flush(a, A);
  A->band = band;
  A->live = 1;
}

struct Ring {
  u32 *qu, *qd, *qe;  // unit (| kFind3Bit), d, e (or c3)
  u32 *qx, *qy;       // strict-gate row words (variant S, survivor ring only)
  u32 head, tail;
  WaveAcc* acc;       // the wavefront's counter sums
};

// probe mode 1: append the hashes of lanes < n to the wavefront's emit chunk
// (a new chunk when this one cannot take them all); without a free chunk,
// query here and keep the accepted hashes
__device__ __forceinline__ void emit_hashes(const BatchedArgs& a, WaveAcc* A, u64 fp, u32 n, u32 lane) {
  const u64 h = mix64(fp + a.main.seed);
  u32 ch = 0, cf = 0;
  if (lane == 0) {
    if (A->chunk == ~0u || (A->chunk < a.n_chunks && A->cfill + n > kEmitChunk)) {
      if (A->chunk != ~0u) a.chunk_fill[A->chunk] = A->cfill;
      A->chunk = atomicAdd(a.chunk_ctr, 1u);
      A->cfill = 0;
    }
    ch = A->chunk;
    cf = A->cfill;
    if (ch < a.n_chunks) A->cfill = cf + n;
  }
  ch = (u32)__shfl((int)ch, 0, 64);
  cf = (u32)__shfl((int)cf, 0, 64);
  if (ch < a.n_chunks) {
    if (lane < n) a.flat[(u64)ch * kEmitChunk + cf + lane] = h;
  } else if (lane < n && d_query(a.main, fp)) {
    const u32 k = atomicAdd(a.acc_n, 1u);
    if (k < a.acc_cap)
      a.acc_h[k] = h;
    else
      atomicAdd(a.acc_lost, 1u);
  }
}

// probe mode 2: is hash h in the accepted set
__device__ __forceinline__ bool in_set(const BatchedArgs& a, u64 h) {
  const u32 bit = (u32)(h >> (64 - kSetFilterBits));
  if (a.set_n == 0 || !((a.set_bits[bit >> 5] >> (bit & 31u)) & 1u)) return false;
  u32 lo = 0, hi = a.set_n;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (a.set_h[mid] < h)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < a.set_n && a.set_h[lo] == h;
}

// probe the ring's oldest n entries (n <= 64, one per lane).  S: also count
// the queries per band (into the wavefront's sums; entries of a second band
// in one batch by an atomic each).
template <bool S>
__device__ __forceinline__ void drain(const BatchedArgs& a, const Ring& q, u32 n, u32 lane) {
  u32 band = 0, uid = 0, d = 0, e = 0;
  u64 fp = 0;
  if (lane < n) {
    const u32 slot = (q.head + lane) & (kRing - 1);
    const u32 tag = q.qu[slot];
    d = q.qd[slot];
    e = q.qe[slot];
    if (tag & kFind3Bit) {  // classes 2-4: V = Q - floor(d^6/mod) - c^6, c in e
      const DevUnit3* U = a.units3 + (tag & ~kFind3Bit);
      const u64 d3 = (u64)d * d * d;
      const u64 dq = (u64)((((u128)d3 * d3) - U->dr) >> U->shift) * U->inv_lo;
      const u64 c2 = (u64)e * e;
      fp = U->Q_lo - dq - c2 * c2 * c2;
      uid = U->uid;
      band = U->band;
    } else {  // class 5: T = Q - floor(e^6/729) - floor(d^6/64)
      const DevUnit5* U = a.units + tag;
      const u64 e2 = (u64)e * e;
      const u64 eq = (e2 * e2 * e2 - (u64)U->te) * a.inv729;
      const u64 d3 = (u64)d * d * d;
      const u64 dq = (u64)(((u128)d3 * d3) >> 6);
      fp = U->Q_lo - eq - dq;
      uid = U->uid;
      band = U->band;
    }
  }
  if (!S && a.pmode == 1u) {
    emit_hashes(a, q.acc, fp, n, lane);
    return;
  }
  if (lane < n) {
    const bool main_hit = a.pmode == 2u ? in_set(a, mix64(fp + a.main.seed)) : d_query(a.main, fp);
    u32 flags = main_hit ? 1u : 0u;
    if (a.side.present && d_query(a.side, fp)) flags |= 2u;
    if (flags) {
      Counters* c = a.ctr + band;
      if (flags & 1u) atomicAdd((unsigned long long*)&c->maybe, 1ULL);
      if (flags & 2u) atomicAdd((unsigned long long*)&c->side_maybe, 1ULL);
      atomicAdd((unsigned long long*)&c->recs, 1ULL);
      const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
      if (s < a.hit_cap) {
        HitRec h;
        h.unit = uid;
        h.d = d;
        h.e = e;
        h.flags = flags | (band << 8);
        h.fp = fp;
        a.hits[s] = h;
      }
    }
  }
  if (S) {
    const u32 b0 = (u32)__builtin_amdgcn_readfirstlane((int)band);  // lane 0 holds an entry
    const u64 same = __ballot(lane < n && band == b0);
    if (lane == 0) {
      acc_band(a, q.acc, b0);
      q.acc->v[7] += (u64)__popcll(same);
    }
    if (lane < n && band != b0) atomicAdd((unsigned long long*)&a.ctr[band].queries, 1ULL);
  }
}

// the ring's LDS writes by some lanes visible to the reads by others
__device__ __forceinline__ void ring_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// variant S: append strict-gate survivors to the probe ring p; probe full batches of 64
__device__ __forceinline__ void push_probe(const BatchedArgs& a, Ring& p, bool pass, u32 tag, u32 x, u32 y,
                                           u32 lane) {
  const u64 m = __ballot(pass);
  if (!m) return;
  if (pass) {
    const u32 slot = (p.tail + lanes_below(m)) & (kRing - 1);
    p.qu[slot] = tag;
    p.qd[slot] = x;
    p.qe[slot] = y;
  }
  p.tail += (u32)__popcll(m);
  if (p.tail - p.head >= 64u) {
    ring_sync();
    drain<true>(a, p, 64u, lane);
    p.head += 64u;
  }
}

// variant S: the strict gate on the survivor ring's oldest n entries (n <= 64);
// T = row part - lane part with L = floor(d^6/64) (class 5) or c3^6 (classes 2-4)
__device__ __forceinline__ void strict_stage(const BatchedArgs& a, const Ring& q, Ring& p, u32 n, u32 lane) {
  bool pass = false;
  u32 tag = 0, x = 0, y = 0;
  if (lane < n) {
    const u32 slot = (q.head + lane) & (kRing - 1);
    tag = q.qu[slot];
    x = q.qd[slot];
    y = q.qe[slot];
    const StrictRow row{q.qx[slot], q.qy[slot]};
    pass = strict_pass(row, (tag & kFind3Bit) ? strict_lane_p6(y) : strict_lane_q64(x));
  }
  push_probe(a, p, pass, tag, x, y, lane);
}

// append the lanes with `pass` set to the survivor ring q; pass full batches
// of 64 to the probes (S: through the strict gate into the probe ring p)
template <bool S>
__device__ __forceinline__ u32 ring_push(const BatchedArgs& a, Ring& q, Ring& p, bool pass, u32 tag, u32 x,
                                         u32 y, StrictRow xw, u32 lane) {
  const u64 m = __ballot(pass);
  if (!m) return 0;
  if (pass) {
    const u32 slot = (q.tail + lanes_below(m)) & (kRing - 1);
    q.qu[slot] = tag;
    q.qd[slot] = x;
    q.qe[slot] = y;
    if (S) {
      q.qx[slot] = xw.w0;
      q.qy[slot] = xw.w1;
    }
  }
  const u32 cnt = (u32)__popcll(m);
  q.tail += cnt;
  if (q.tail - q.head >= 64u) {
    ring_sync();
    if (S) strict_stage(a, q, p, 64u, lane);
    else drain<false>(a, q, 64u, lane);
    q.head += 64u;
  }
  return cnt;
}

// an item's counters (per-row values on lanes 0..E-1; probes uniform) into the wavefront's sums
template <int E>
__device__ __forceinline__ void flush_counters(const BatchedArgs& a, WaveAcc* A, u32 band, u32 lane, u32 calls,
                                               u32 neg, u32 zero, u32 pos, u32 probes, u32 jobs, u32 failT) {
  // lanes E.. hold zeros; an item's sums stay below 2^32 (at most 32 rows of 2^26 values)
  const u64 s_calls = wave_sum_u32(calls), s_neg = wave_sum_u32(neg), s_zero = wave_sum_u32(zero);
  const u64 s_pos = wave_sum_u32(pos), s_jobs = wave_sum_u32(jobs), s_fail = wave_sum_u32(failT);
  if (lane == 0) {
    acc_band(a, A, band);
    A->v[0] += s_calls;
    A->v[1] += s_neg;
    A->v[2] += s_zero;
    A->v[3] += s_pos;
    A->v[4] += probes;
    A->v[5] += s_jobs;
    A->v[6] += s_fail;
  }
}

// one class-5 item: E e-rows of one unit, every d
template <int E, bool S, bool V>
__device__ __forceinline__ void cls5_item(const BatchedArgs& a, u32 u0, u32 n_units, const u32* s_gate,
                                          const u32* s_sieve, Ring& q, Ring& p, u32 item, u32 lane) {
  const u32 K = a.ndg_K;
  u32 lo = u0, hi = u0 + n_units;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (a.units[mid].item_base <= item) lo = mid;
    else hi = mid;
  }
  const u32 ui = lo;
  const DevUnit5 U = a.units[ui];
  const u32 row0 = (item - U.item_base) * (u32)E;
  const u32 r0 = U.res2[0], r1 = U.res2[1], r2 = U.res2[2], r3 = U.res2[3];
  const u32 minr = min(min(r0, r1), min(r2, r3));

  // per-row set-up on lanes 0..E-1: e, T-space row value, positive cut
  u32 e = 0, dlim = 0, xg = 0, sa = 0, sb = 0, sc = 0;
  StrictRow xs{0u, 0u};
  StrictSieveRow st{};
  u32 c_calls = 0, c_neg = 0, c_zero = 0, c_pos = 0;
  if (lane < (u32)E) {
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
          xg = mod_gate(X);
          if (V) sieve_row_bases(xg, U.td, sa, sb, sc);
          if (S) xs = strict_row((u64)(X >> 64), (u64)X);
          if (V && S) st = strict_sieve_bases(xs, U.td);
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
  const u32 maxdl = wave_max_u32(dlim);

  u32 probes_item = 0;
  if (V) {
    // bit sieve: lane slot = (word w, residue index) over the item's d range,
    // one word = 32 consecutive k of one residue; per row five LDS words AND-ed,
    // the cut applied, survivors taken lowest bit first.  S: five strict words
    // more (global memory), and the survivors go straight to the probe ring.
    const u32 nslots = maxdl >= minr ? 4u * ((maxdl - minr) / 2048u + 1u) : 0u;
    for (u32 s0 = 0; s0 < nslots; s0 += 64u) {
      const u32 slot = s0 + lane, ri = slot & 3u, w = slot >> 2;
      const u32 r = ri == 0u ? r0 : ri == 1u ? r1 : ri == 2u ? r2 : r3;
      const u32 d0 = slot < nslots ? r + 2048u * w : 0xFFFFFFFFu;  // past the slots: no grid point
      const u32 j9 = d0 % 3u, j7 = d0 % 7u, j13 = d0 % 13u, j19 = d0 % 19u;
      StrictSieveLane sj{};
      if (S) sj = strict_sieve_lane(d0);
      const u32 dminp = minr + 2048u * (s0 >> 2);  // the pass's least d0
      for (int i = 0; i < E; ++i) {
        const u32 dl = rdl(dlim, i);
        if (dl < dminp) continue;
        u32 P = sieve_cut(d0, dl);
        if (a.use_gate) {
          const u32 ra = rdl(sa, i), rb = rdl(sb, i), rc = rdl(sc, i);
          P &= s_sieve[(ra & 0xFFFFu) + r] & s_sieve[(ra >> 16) + j9] & s_sieve[(rb & 0xFFFFu) + j7] &
               s_sieve[(rb >> 16) + j13] & s_sieve[rc + j19];
        }
        probes_item += (u32)__popc(P);
        if (S) {
          StrictSieveRow b;
#pragma unroll
          for (int k = 0; k < kStrictRowWords; ++k) b.w[k] = rdl(st.w[k], i);
          if (P) P &= strict_sieve_word(g_strict_sieve, b, sj, r);
        }
        const u32 ei = rdl(e, i);
        while (__ballot(P != 0u)) {
          const bool has = P != 0u;
          const u32 d = d0 + 64u * (has ? (u32)__builtin_ctz(P) : 0u);
          if (S) push_probe(a, p, has, ui, d, ei, lane);
          else ring_push<false>(a, q, p, has, ui, d, ei, StrictRow{0u, 0u}, lane);
          P &= P - 1u;
        }
      }
    }
    flush_counters<E>(a, q.acc, U.band, lane, c_calls, c_neg, c_zero, c_pos, wave_sum_u32(probes_item), 0u, 0u);
    return;
  }
  // gate over the positive part of the rows; lanes run over k
  for (u32 kb = 0; minr + 64u * kb <= maxdl; kb += 64u) {
    const u32 k = kb + lane;
    const u32 d0 = r0 + 64u * k, d1 = r1 + 64u * k, d2 = r2 + 64u * k, d3 = r3 + 64u * k;
    const u32 n0 = d0 <= U.fmax2 ? a.ndg[r0 * K + k] : 0u;
    const u32 n1 = d1 <= U.fmax2 ? a.ndg[r1 * K + k] : 0u;
    const u32 n2 = d2 <= U.fmax2 ? a.ndg[r2 * K + k] : 0u;
    const u32 n3 = d3 <= U.fmax2 ? a.ndg[r3 * K + k] : 0u;
    const u32 dmin_blk = minr + 64u * kb;
    for (int i = 0; i < E; ++i) {
      const u32 dl = rdl(dlim, i);
      if (dl < dmin_blk) continue;
      const u32 xgi = rdl(xg, i);
      const u32 ei = rdl(e, i);
      const StrictRow xsi = S ? StrictRow{rdl(xs.w0, i), rdl(xs.w1, i)} : StrictRow{0u, 0u};
      const u32 ds[4] = {d0, d1, d2, d3};
      const u32 ns[4] = {n0, n1, n2, n3};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool pass = gate_pass(s_gate, xgi, ns[j]) && ds[j] <= dl;
        probes_item += ring_push<S>(a, q, p, pass, ui, ds[j], ei, xsi, lane);
      }
    }
  }
  flush_counters<E>(a, q.acc, U.band, lane, c_calls, c_neg, c_zero, c_pos, probes_item, 0u, 0u);
}

// one class 2-4 item: E d-rows of one unit, every c3 of each row's window
// (fourcore_find_v4.cu:165-205 expansion, :401-425 window, :255-284 sweep)
template <int E, bool S, bool V>
__device__ __forceinline__ void find3_item(const BatchedArgs& a, u32 u0, u32 n_units, const u32* s_gate,
                                           const u32* s_csieve, Ring& q, Ring& p, u32 item, u32 lane) {
  u32 lo = u0, hi = u0 + n_units;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (a.units3[mid].item_base <= item) lo = mid;
    else hi = mid;
  }
  const u32 ui = lo;
  const DevUnit3* Up = a.units3 + ui;
  const u32 row0 = (item - Up->item_base) * (u32)E;
  u32 d = 0, wlo = 1, whi = 0, xg = 0, ca = 0, cb = 0, cc = 0;
  StrictRow xs{0u, 0u};
  StrictSieveRow cs{};
  u32 c_calls = 0, c_jobs = 0, c_fail = 0;
  if (lane < (u32)E) {
    const u32 r = row0 + lane;
    if (r < Up->rows) {
      d = (u32)Up->res[r % Up->nres] + (r / Up->nres) * Up->mod;
      if (d >= 1 && d <= Up->fmax) {
        const u64 d3 = (u64)d * d * d;
        const u128 inv = ((u128)Up->inv_hi << 64) | Up->inv_lo;
        const u128 dq = ((((u128)d3 * d3) - Up->dr) >> Up->shift) * inv;  // exact floor(d^6/mod)
        const u128 Q = ((u128)Up->Q_hi << 64) | Up->Q_lo;
        if (Q > dq) {
          const u128 X = Q - dq;  // the job's T
          c_jobs = 1;
          u32 h = iroot6_u128(X);
          if (h > Up->lim) h = Up->lim;
          const u32 l = min_core(X, 3u);  // smallest c >= 1 with 3 c^6 >= T
          if (l <= h) {
            wlo = l;
            whi = h;
            c_calls = h - l + 1;
          }
          xg = mod_gate(X);
          if (V) csieve_row_bases(xg, ca, cb, cc);
          if (S) xs = strict_row((u64)(X >> 64), (u64)X);
          if (V && S) cs = strict_csieve_bases(xs);
        } else {
          c_fail = 1;  // not positive: compute_T fails (fourcore_gmp.hpp:47)
        }
      }
    }
  }
  u32 probes_item = 0;
  if (V) {
    // bit sieve over c: a lane takes 32 consecutive c of the row's window, the
    // five reference words (LDS) and, S, the ten strict words (global memory)
    // AND-ed; the survivors (S: strict survivors, straight to the probe ring)
    // taken lowest bit first
    for (int i = 0; i < E; ++i) {
      const u32 l = rdl(wlo, i), h = rdl(whi, i);
      if (l > h) continue;
      const u32 di = rdl(d, i), ra = rdl(ca, i), rb = rdl(cb, i), rc = rdl(cc, i);
      StrictSieveRow sb{};
      if (S) {
#pragma unroll
        for (int k = 0; k < kStrictRowWords; ++k) sb.w[k] = rdl(cs.w[k], i);
      }
      const u32 nw = (h - l) / 32u + 1u;
      for (u32 w0 = 0; w0 < nw; w0 += 64u) {
        const u32 w = w0 + lane;
        const u32 c0 = l + 32u * w;  // below 2^22: l, h < 2^21
        u32 P = 0;
        if (w < nw) {
          P = csieve_cut(c0, h);
          if (a.use_gate) P &= csieve_word(s_csieve, ra, rb, rc, c0);
        }
        probes_item += (u32)__popc(P);
        if (S && P) P &= strict_csieve_word(g_strict_csieve, sb, c0);
        while (__ballot(P != 0u)) {
          const bool has = P != 0u;
          const u32 c = c0 + (has ? (u32)__builtin_ctz(P) : 0u);
          if (S) push_probe(a, p, has, ui | kFind3Bit, di, c, lane);
          else ring_push<false>(a, q, p, has, ui | kFind3Bit, di, c, StrictRow{0u, 0u}, lane);
          P &= P - 1u;
        }
      }
    }
    flush_counters<E>(a, q.acc, Up->band, lane, c_calls, 0u, 0u, c_calls, wave_sum_u32(probes_item), c_jobs, c_fail);
    return;
  }
  for (int i = 0; i < E; ++i) {
    const u32 l = rdl(wlo, i), h = rdl(whi, i);
    if (l > h) continue;
    const u32 xgi = rdl(xg, i), di = rdl(d, i);
    const StrictRow xsi = S ? StrictRow{rdl(xs.w0, i), rdl(xs.w1, i)} : StrictRow{0u, 0u};
    for (u32 cb = l; cb <= h; cb += 64u) {
      const u32 c = cb + lane;
      const bool in = c <= h;
      const u32 n = in ? neg_p6_gate(c) : 0u;  // c <= lim < 2^21
      const bool pass = in && gate_pass(s_gate, xgi, n);
      probes_item += ring_push<S>(a, q, p, pass, ui | kFind3Bit, di, c, xsi, lane);
      if (h - cb < 64u) break;  // cb + 64 could wrap past 2^32
    }
  }
  flush_counters<E>(a, q.acc, Up->band, lane, c_calls, 0u, 0u, c_calls, probes_item, c_jobs, c_fail);
}

template <int E, bool S, bool V>
__global__ __launch_bounds__(256) void k_cls5_batched(BatchedArgs a) {
  constexpr u32 kSW = S ? kWaves : 1, kSR = S ? kRing : 1;  // variant S only: row words, probe ring
  __shared__ u32 s_gate[kGateWords];
  __shared__ u32 s_sieve[V ? kSieveWords : 1], s_csieve[V ? kCSieveWords : 1];
  __shared__ u32 s_qu[kWaves][kRing], s_qd[kWaves][kRing], s_qe[kWaves][kRing];
  __shared__ u32 s_qx[kSW][kSR], s_qy[kSW][kSR], s_pu[kSW][kSR], s_pd[kSW][kSR], s_pe[kSW][kSR];
  __shared__ WaveAcc s_acc[kWaves];
  for (u32 i = threadIdx.x; i < kGateWords; i += blockDim.x) s_gate[i] = a.gate[i];
  if (V) {
    for (u32 i = threadIdx.x; i < kSieveWords; i += blockDim.x) s_sieve[i] = kSieveDev.w[i];
    for (u32 i = threadIdx.x; i < kCSieveWords; i += blockDim.x) s_csieve[i] = kCSieveDev.w[i];
  }
  __syncthreads();

  const u32 lane = threadIdx.x & 63u;
  const u32 w = threadIdx.x >> 6;
  const u32 ws = S ? w : 0u;
  WaveAcc* acc = &s_acc[w];
  if (lane == 0) {
    acc->live = 0;
    acc->chunk = ~0u;
    acc->cfill = 0;
    for (int i = 0; i < 8; ++i) acc->v[i] = 0;
  }
  Ring q{s_qu[w], s_qd[w], s_qe[w], s_qx[ws], s_qy[ws], 0u, 0u, acc};
  Ring p{s_pu[ws], s_pd[ws], s_pe[ws], nullptr, nullptr, 0u, 0u, acc};
  const u32 grab = a.grab ? a.grab : 1u;  // consecutive items per work-counter update
  for (;;) {
    u32 item0 = 0;
    // probe mode 1: once this wavefront's latest chunk (handed out in order)
    // is past the limit, the emit chunks are nearly used up: the work items
    // left go to the next launch
    if (lane == 0)
      item0 = (a.pmode == 1u && acc->chunk != ~0u && acc->chunk >= a.chunk_limit) ? ~0u : atomicAdd(a.item_ctr, grab);
    item0 = __shfl(item0, 0, 64);
    if (item0 >= a.total) break;
    const u32 iend = a.total - item0 > grab ? item0 + grab : a.total;
    for (u32 item = item0; item < iend; ++item) {
      u32 lo = 0, hi = a.n_segs;  // the segment holding this launch item
      while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (a.segs[mid].g0 <= item) lo = mid;
        else hi = mid;
      }
      const Segment G = a.segs[lo];
      const u32 local = G.item0 + (item - G.g0);
      if (G.kind == 0) cls5_item<E, S, V>(a, G.u0, G.n_units, s_gate, s_sieve, q, p, local, lane);
      else find3_item<E, S, V>(a, G.u0, G.n_units, s_gate, s_csieve, q, p, local, lane);
    }
  }
  if (q.tail != q.head) {
    ring_sync();
    if (S) strict_stage(a, q, p, q.tail - q.head, lane);
    else drain<false>(a, q, q.tail - q.head, lane);
  }
  if (S) {
    while (p.tail != p.head) {
      const u32 n = min(64u, p.tail - p.head);
      ring_sync();
      drain<true>(a, p, n, lane);
      p.head += n;
    }
  }
  if (lane == 0) {
    acc_flush(a, acc);
    if (a.pmode == 1u && acc->chunk < a.n_chunks) a.chunk_fill[acc->chunk] = acc->cfill;
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

// T mod m from R = 42^6 T: (R mod 42^6 m) / 42^6 (strict gate of the plain kernel)
constexpr u64 kStrictA42 = (u64)kStrictA * kM42, kStrict81_42 = 81ULL * kM42;
constexpr u64 kStrictB42 = (u64)kStrictB * kM42, kStrictC42 = (u64)kStrictC * kM42;
constexpr u64 kTwo64A42 = (u64)(((u128)1 << 64) % kStrictA42), kTwo64_81_42 = (u64)(((u128)1 << 64) % kStrict81_42);
constexpr u64 kTwo64B42 = (u64)(((u128)1 << 64) % kStrictB42), kTwo64C42 = (u64)(((u128)1 << 64) % kStrictC42);

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
  u64 calls = 0, neg = 0, zero = 0, gated = 0, probes = 0, queries = 0;
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
    if (a.strict) {
      const u32 ta = (u32)(mod192(x2, x1, x0, kStrictA42, kTwo64A42) / kM42);
      const u32 t81 = (u32)(mod192(x2, x1, x0, kStrict81_42, kTwo64_81_42) / kM42);
      const u32 tb = (u32)(mod192(x2, x1, x0, kStrictB42, kTwo64B42) / kM42);
      const u32 tc = (u32)(mod192(x2, x1, x0, kStrictC42, kTwo64C42) / kM42);
      if (!strict_pass_t(ta, t81, (u32)fp & 63u, tb, tc)) continue;
      ++queries;
    }
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
  queries = wave_sum(queries);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd((unsigned long long*)&a.ctr->calls, (unsigned long long)calls);
    atomicAdd((unsigned long long*)&a.ctr->neg, (unsigned long long)neg);
    atomicAdd((unsigned long long*)&a.ctr->zero, (unsigned long long)zero);
    atomicAdd((unsigned long long*)&a.ctr->gated, (unsigned long long)gated);
    atomicAdd((unsigned long long*)&a.ctr->probes, (unsigned long long)probes);
    if (queries) atomicAdd((unsigned long long*)&a.ctr->queries, (unsigned long long)queries);
  }
}

// ---------------------------------------------------------------------------
// plain find3 (the reference kernel shape, fourcore_find_v4.cu:255-284):
// block per candidate, threads stride c3 over [lo, hi]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_find3_plain(PlainF3Args a) {
  const int ci = blockIdx.x;
  if (ci >= a.n) return;
  const PlainCand3 C = a.cands[ci];
  __shared__ u16 s504[504], s247[247];
  __shared__ u64 g504[8], g247[4];
  for (u32 k = threadIdx.x; k < 504; k += blockDim.x) s504[k] = a.p6[k];
  for (u32 k = threadIdx.x; k < 247; k += blockDim.x) s247[k] = a.p6[504 + k];
  if (threadIdx.x < 8) g504[threadIdx.x] = a.gate.g504[threadIdx.x];
  if (threadIdx.x < 4) g247[threadIdx.x] = a.gate.g247[threadIdx.x];
  __syncthreads();
  u64 calls = 0, gated = 0, probes = 0, queries = 0;
  for (u32 c3 = C.lo + threadIdx.x; c3 <= C.hi && c3 >= C.lo; c3 += blockDim.x) {
    ++calls;
    if (a.use_gate) {
      u32 r504 = C.r504 + 504u - s504[c3 % 504];
      if (r504 >= 504u) r504 -= 504u;
      u32 r247 = C.r247 + 247u - s247[c3 % 247];
      if (r247 >= 247u) r247 -= 247u;
      if (!(((g504[r504 >> 6] >> (r504 & 63)) & 1ULL) && ((g247[r247 >> 6] >> (r247 & 63)) & 1ULL))) {
        ++gated;
        continue;
      }
    }
    ++probes;
    if (a.strict) {
      if (!strict_pass(StrictRow{C.srow, C.srow1}, strict_lane_p6(c3))) continue;
      ++queries;
    }
    const u64 c2 = (u64)c3 * c3;
    const u64 fp = C.q_lo - c2 * c2 * c2;
    u32 flags = d_query(a.main, fp) ? 1u : 0u;
    if (a.side.present && d_query(a.side, fp)) flags |= 2u;
    if (flags) {
      Counters* c = a.ctr + C.band;
      if (flags & 1u) atomicAdd((unsigned long long*)&c->maybe, 1ULL);
      if (flags & 2u) atomicAdd((unsigned long long*)&c->side_maybe, 1ULL);
      atomicAdd((unsigned long long*)&c->recs, 1ULL);
      const u64 s = atomicAdd((unsigned long long*)a.nhits, 1ULL);
      if (s < a.hit_cap) {
        HitRec h;
        h.unit = C.uid;
        h.d = C.d;
        h.e = c3;
        h.flags = flags;
        h.fp = fp;
        a.hits[s] = h;
      }
    }
  }
  calls = wave_sum(calls);
  gated = wave_sum(gated);
  probes = wave_sum(probes);
  queries = wave_sum(queries);
  if ((threadIdx.x & 63u) == 0) {
    Counters* c = a.ctr + C.band;
    atomicAdd((unsigned long long*)&c->calls, (unsigned long long)calls);
    atomicAdd((unsigned long long*)&c->gated, (unsigned long long)gated);
    atomicAdd((unsigned long long*)&c->probes, (unsigned long long)probes);
    if (queries) atomicAdd((unsigned long long*)&c->queries, (unsigned long long)queries);
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

template <bool S, bool V>
static int launch_batched_variant(const BatchedArgs& a, int grid, int rows, hipStream_t s) {
  switch (rows) {
    case 4: hipLaunchKernelGGL((k_cls5_batched<4, S, V>), dim3(grid), dim3(64 * kWaves), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_cls5_batched<8, S, V>), dim3(grid), dim3(64 * kWaves), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_cls5_batched<16, S, V>), dim3(grid), dim3(64 * kWaves), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_cls5_batched<32, S, V>), dim3(grid), dim3(64 * kWaves), 0, s, a); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int upload_sieve_tables() {
  std::vector<u32> w(kStrictWords), c(kStrictCWords);
  fill_strict_sieve(w.data());
  fill_strict_csieve(c.data());
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_strict_sieve), w.data(), kStrictWords * sizeof(u32), 0,
                                   hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpyToSymbol(HIP_SYMBOL(g_strict_csieve), c.data(), kStrictCWords * sizeof(u32), 0,
                          hipMemcpyHostToDevice);
  return (int)e;
}

const u32* strict_csieve_table() {
  void* p = nullptr;
  return hipGetSymbolAddress(&p, HIP_SYMBOL(g_strict_csieve)) == hipSuccess ? static_cast<const u32*>(p) : nullptr;
}

int launch_cls5_batched(const BatchedArgs& args, int grid, int rows, hipStream_t s) {
  // items per work-counter update: 4 when every wavefront has at least 256
  // items to take, fewer for short launches (the tail stays balanced)
  BatchedArgs a = args;
  const u64 per_wave = (u64)a.total / ((u64)(grid > 0 ? grid : 1) * kWaves);
  a.grab = per_wave >= 256 ? 4u : per_wave >= 64 ? 2u : 1u;
  if (a.pmode > 2u || (a.pmode && a.strict)) return (int)hipErrorInvalidValue;
  if (a.strict) return a.sieve ? launch_batched_variant<true, true>(a, grid, rows, s)
                               : launch_batched_variant<true, false>(a, grid, rows, s);
  return a.sieve ? launch_batched_variant<false, true>(a, grid, rows, s)
                 : launch_batched_variant<false, false>(a, grid, rows, s);
}

int launch_cls5_plain(const PlainArgs& a, int n_units, int ymax, hipStream_t s) {
  hipLaunchKernelGGL(k_cls5_plain, dim3(n_units, ymax), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

int launch_find3_plain(const PlainF3Args& a, hipStream_t s) {
  hipLaunchKernelGGL(k_find3_plain, dim3(a.n), dim3(256), 0, s, a);
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

template <bool S, bool V>
static hipError_t occupancy_variant(int& nb, int rows) {
  switch (rows) {
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cls5_batched<4, S, V>, 64 * kWaves, 0);
    case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cls5_batched<8, S, V>, 64 * kWaves, 0);
    case 16: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cls5_batched<16, S, V>, 64 * kWaves, 0);
    case 32: return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cls5_batched<32, S, V>, 64 * kWaves, 0);
    default: return hipErrorInvalidValue;
  }
}

int batched_blocks_per_cu(int rows, bool strict, bool sieve) {
  int nb = 0;
  const hipError_t e = strict ? (sieve ? occupancy_variant<true, true>(nb, rows) : occupancy_variant<true, false>(nb, rows))
                              : (sieve ? occupancy_variant<false, true>(nb, rows) : occupancy_variant<false, false>(nb, rows));
  if (e != hipSuccess) return 4;
  return nb > 0 ? nb : 1;
}

}  // namespace fc5
