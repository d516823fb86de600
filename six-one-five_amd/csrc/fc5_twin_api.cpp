// fc5_twin_api.cpp -- C ABI of the (6,1,7) class-5 twin search (include/fc5.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "fc5_ctx.hpp"

using namespace fc5;

namespace {

u64 inv_mod_small(u64 a, u64 m) {  // a coprime to m, m small
  for (u64 x = 1; x < m; ++x)
    if (a * x % m == 1) return x;
  return 0;
}

// triple-sum sets of sixth powers mod 504 and 247, 0 included
// (build_tri_set, solve_617_v1.cu:334-346): words 0..15 and 16..23
std::vector<u32> tri_words() {
  std::vector<u32> w(24, 0);
  const u32 ms[2] = {504, 247};
  for (int k = 0; k < 2; ++k) {
    const u32 m = ms[k];
    std::vector<char> p1(m, 0), p2(m, 0), p3(m, 0);
    for (u32 i = 0; i < m; ++i) p1[(u32)pow6_mod(i, m)] = 1;
    for (u32 a = 0; a < m; ++a)
      if (p1[a])
        for (u32 c = 0; c < m; ++c)
          if (p1[c]) p2[(a + c) % m] = 1;
    for (u32 a = 0; a < m; ++a)
      if (p2[a])
        for (u32 c = 0; c < m; ++c)
          if (p1[c]) p3[(a + c) % m] = 1;
    for (u32 x = 0; x < m; ++x)
      if (p3[x]) w[16 * k + (x >> 5)] |= 1u << (x & 31);
  }
  return w;
}

TwinCandRaw raw_of(const fc5_twin_cand& c) {
  TwinCandRaw r{};
  r.B = c.B;
  r.u = c.u;
  r.w = c.w;
  r.base = ((u128)c.base_hi << 64) | c.base_lo;
  r.lim = c.lim;
  r.fmax1 = c.fmax1;
  r.mod1 = c.mod1;
  r.nres1 = c.nres1;
  std::memcpy(r.res, c.res, sizeof r.res);
  return r;
}

fc5_twin_cand abi_of(const TwinCandRaw& r) {
  fc5_twin_cand c{};
  c.B = r.B;
  c.u = r.u;
  c.w = r.w;
  c.base_lo = (u64)r.base;
  c.base_hi = (u64)(r.base >> 64);
  c.lim = r.lim;
  c.fmax1 = r.fmax1;
  c.mod1 = r.mod1;
  c.nres1 = r.nres1;
  std::memcpy(c.res, r.res, sizeof c.res);
  return c;
}

int ensure_twin(fc5_ctx* ctx, size_t n) {
  if (!ctx->d_tri) {
    const auto w = tri_words();
    if (hipMalloc(&ctx->d_tri, 24 * 4) != hipSuccess || hipMalloc(&ctx->d_twin_ctr, sizeof(TwinCounters)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_twin_ctr, sizeof(TwinCounters), hipHostMallocDefault) != hipSuccess)
      FAIL(ctx, FC5_E_NOMEM, "twin buffers");
    HIPCHK(ctx, hipMemcpy(ctx->d_tri, w.data(), 24 * 4, hipMemcpyHostToDevice));
  }
  if (n > ctx->d_twin_cap) {
    if (ctx->d_twin) (void)hipFree(ctx->d_twin);
    ctx->d_twin = nullptr;
    ctx->d_twin_cap = 0;
    const size_t cap = n + n / 4 + 64;
    if (hipMalloc(&ctx->d_twin, cap * sizeof(DevTwinCand)) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "twin candidates");
    ctx->d_twin_cap = cap;
  }
  return FC5_OK;
}

// one persistent launch over packed candidates; appends hits to ctx->twin_hits
// (hit candidate indices are the uid values)
int twin_launch(fc5_ctx* ctx, const std::vector<TwinCand>& cands, const std::vector<u32>& uid,
                fc5_twin_result* res) {
  std::vector<DevTwinCand> dev(cands.size());
  u64 items = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    const TwinCand& C = cands[i];
    DevTwinCand& D = dev[i];
    std::memset(&D, 0, sizeof D);
    D.Q_lo = (u64)C.Q;
    D.Q_hi = (u64)(C.Q >> 64);
    D.fmax1 = C.fmax1;
    D.lim = C.lim;
    D.rows = C.nres * (C.fmax1 / 64 + 1);
    D.item_base = (u32)items;
    D.uid = uid[i];
    D.nres = C.nres;
    for (u32 k = 0; k < C.nres; ++k) D.res[k] = (u8)C.res[k];
    items += D.rows;
  }
  if (items > kMaxItems) FAIL(ctx, FC5_E_ARG, "too many twin work items in one launch (more than 2^32 - 2^20)");
  if (ctx->strict_gate && ctx->twin_plain)
    FAIL(ctx, FC5_E_ARG, "the strict gate runs in the batched twin kernel only (twin_plain = 1 with strict_gate = 1)");
  int rc = ensure_twin(ctx, dev.size());
  if (rc != FC5_OK) return rc;
  if ((rc = ensure_hits(ctx)) != FC5_OK) return rc;
  if (!dev.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_twin, dev.data(), dev.size() * sizeof(DevTwinCand), hipMemcpyHostToDevice));
  TwinArgs a{};
  a.cands = ctx->d_twin;
  a.n_cands = (u32)dev.size();
  a.n_items = (u32)items;
  a.gate = ctx->d_gate;
  a.tri = ctx->d_tri;
  a.use_gate = (u32)ctx->use_gate;
  a.use_tri = (u32)ctx->twin_tri;
  a.c4_div = (u32)ctx->twin_c4_div;
  a.tile = (u32)ctx->twin_tile;
  a.strict = (u32)ctx->strict_gate;
  a.scsieve = strict_csieve_table();
  if (a.strict && !a.scsieve) FAIL(ctx, FC5_E_HIP, "strict sieve table not on the device");
  a.main = ctx->main.dev;
  a.hits = reinterpret_cast<TwinHit*>(ctx->d_hits);
  a.hit_cap = ctx->hit_cap;
  a.nhits = ctx->d_nhits;
  a.ctr = ctx->d_twin_ctr;
  a.item_ctr = ctx->d_item_ctr;
  const u64 zero_words = sizeof(TwinCounters) / 8;
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (launch_fill_u64(reinterpret_cast<u64*>(ctx->d_twin_ctr), zero_words, 0, ctx->stream) != 0 ||
      launch_fill_u64(ctx->d_nhits, 1, 0, ctx->stream) != 0 ||
      launch_fill_u64(reinterpret_cast<u64*>(ctx->d_item_ctr), 1, 0, ctx->stream) != 0)
    FAIL(ctx, FC5_E_HIP, "counter reset");
  std::vector<PlainTwinCand> plain;
  PlainTwinCand* d_plain = nullptr;
  struct FreeOnExit {  // hipFree waits for the launch that reads the buffer
    PlainTwinCand*& p;
    ~FreeOnExit() {
      if (p) (void)hipFree(p);
    }
  } free_plain{d_plain};
  if (items && ctx->twin_plain) {
    // the reference launch shape over the same candidates
    u32 ymax = 1, lmax = 0;
    plain.resize(cands.size());
    for (size_t i = 0; i < cands.size(); ++i) {
      const TwinCand& C = cands[i];
      PlainTwinCand& P = plain[i];
      std::memset(&P, 0, sizeof P);
      P.b_lo = (u64)C.base;
      P.b_hi = (u64)(C.base >> 64);
      P.lim = C.lim;
      P.fmax1 = C.fmax1;
      P.nres = C.nres;
      P.uid = uid[i];
      for (u32 k = 0; k < C.nres; ++k) P.res[k] = C.res[k];
      ymax = std::max(ymax, C.nres * (C.fmax1 / 64 + 1));
      lmax = std::max(lmax, C.lim);
    }
    if (hipMalloc(&d_plain, plain.size() * sizeof(PlainTwinCand)) != hipSuccess)
      FAIL(ctx, FC5_E_NOMEM, "plain twin candidates");
    PlainTwinArgs p{};
    p.cands = d_plain;
    p.n_cands = (u32)plain.size();
    p.use_gate = (u32)ctx->use_gate;
    p.use_tri = (u32)ctx->twin_tri;
    p.c4_div = (u32)ctx->twin_c4_div;
    p.gate = ctx->gw;
    p.tri = ctx->d_tri;
    p.p6 = ctx->d_p6;
    p.inv216 = inv_mod_2_64(85766121ULL);
    p.c_m504 = (u64)(((u128)1 << 64) % (u128)(504ULL * kM42));
    p.c_m247 = (u64)(((u128)1 << 64) % (u128)247ULL);
    p.inv42_247 = (u64)inv_mod_small((u64)pow6_mod(42, 247), 247);
    p.main = ctx->main.dev;
    p.hits = a.hits;
    p.hit_cap = a.hit_cap;
    p.nhits = a.nhits;
    p.ctr = a.ctr;
    // bounded 2048-wide tiles, or the reference launch's zmax = ceil(max lim / tile)
    // unbounded tiles, capped at 32 (solve_617_v1.cu:961-979)
    const u32 zmax = ctx->twin_tile ? std::min<u32>(32, std::max<u32>(1, (lmax + ctx->twin_tile - 1) / ctx->twin_tile))
                                    : lmax / 2048 + 1;
    p.tile = (u32)ctx->twin_tile;
    if (hipMemcpy(d_plain, plain.data(), plain.size() * sizeof(PlainTwinCand), hipMemcpyHostToDevice) != hipSuccess ||
        launch_twin_plain(p, ymax, zmax, ctx->stream) != 0)
      FAIL(ctx, FC5_E_HIP, "plain twin launch");
  } else if (items) {
    const int bpc = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : twin_blocks_per_cu(ctx->strict_gate != 0);
    const u64 want = (items + 3) / 4;  // 4 waves per workgroup
    const int grid = (int)std::max<u64>(1, std::min<u64>((u64)ctx->num_cu * (u64)bpc, want));
    if (launch_twin_batched(a, grid, ctx->stream) != 0) FAIL(ctx, FC5_E_HIP, "twin launch");
  }
  HIPCHK(ctx, hipMemcpyAsync(ctx->h_twin_ctr, ctx->d_twin_ctr, sizeof(TwinCounters), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->h_nhits, ctx->d_nhits, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const TwinCounters c = *ctx->h_twin_ctr;
  const u64 nh = *ctx->h_nhits;
  const u64 stored = std::min<u64>(nh, ctx->hit_cap);
  std::vector<TwinHit> raw(stored);
  if (stored)
    HIPCHK(ctx, hipMemcpy(raw.data(), ctx->d_hits, stored * sizeof(TwinHit), hipMemcpyDeviceToHost));
  for (const TwinHit& h : raw) ctx->twin_hits.push_back(fc5_twin_hit{h.cand, h.d, h.c4, h.c3, h.fp});
  res->jobs += c.jobs;
  res->c4_total += c.c4_total;
  res->c4_tri_skipped += c.c4_tri_skipped;
  res->calls += c.calls;
  res->gated += c.gated;
  res->probes += 3 * c.calls;
  res->maybe += c.maybe;
  if (ctx->strict_gate) res->strict_out += c.calls - c.queries;
  res->n_hits += nh;
  res->hits_dropped += nh - stored;
  res->cands += cands.size();
  res->items += items;
  res->kernel_ms += ms;
  return FC5_OK;
}

// verification of the collected hits against the packed candidates
int twin_verify(fc5_ctx* ctx, const std::vector<TwinCand>& all) {
  std::sort(ctx->twin_hits.begin(), ctx->twin_hits.end(), [](const fc5_twin_hit& x, const fc5_twin_hit& y) {
    if (x.cand != y.cand) return x.cand < y.cand;
    if (x.d != y.d) return x.d < y.d;
    if (x.c4 != y.c4) return x.c4 < y.c4;
    return x.c3 < y.c3;
  });
  if (!ctx->verify || ctx->twin_hits.empty()) return FC5_OK;
  const int N = ctx->main.hdr.N ? (int)ctx->main.hdr.N : 0;
  std::vector<TwinHitIn> in;
  in.reserve(ctx->twin_hits.size());
  for (const fc5_twin_hit& h : ctx->twin_hits) in.push_back(TwinHitIn{h.cand, h.d, h.c4, h.c3, h.fp});
  std::vector<PairTarget> targets;
  twin_targets(N, all, in, targets);
  std::vector<std::vector<std::pair<u32, u32>>> pairs;
  const int rc = device_pairs(ctx, targets, pairs);
  if (rc != FC5_OK) return rc;
  std::vector<std::vector<TwinSolution>> per(in.size());
  for (size_t i = 0; i < in.size(); ++i)
    for (auto [a, b] : pairs[i]) twin_finish(all[in[i].cand], in[i].d, in[i].c4, in[i].c3, a, b, per[i]);
  std::set<std::array<i64, 7>> reported;  // distinct tuples only (solve_617_v1.cu:1096)
  for (const auto& v : per) {
    for (const TwinSolution& S : v) {
      std::array<i64, 7> key;
      std::copy(S.terms, S.terms + 7, key.begin());
      if (!reported.insert(key).second) continue;
      fc5_twin_solution o{};
      o.B = S.B;
      o.u = S.u;
      o.w = S.w;
      o.d = S.d;
      o.c4 = S.c4;
      o.c3 = S.c3;
      o.a = S.a;
      o.b = S.b;
      std::copy(S.terms, S.terms + 7, o.terms);
      o.gcd = S.gcd;
      ctx->twin_sols.push_back(o);
    }
    if (ctx->stop_first && !ctx->twin_sols.empty()) break;
  }
  return FC5_OK;
}

int finish(fc5_ctx* ctx, fc5_twin_result* res, std::chrono::steady_clock::time_point t0) {
  res->solutions = ctx->twin_sols.size();
  res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (res->hits_dropped) {
    ctx->err = "hit buffer overflow: " + std::to_string(res->hits_dropped) + " hits dropped";
    return FC5_E_OVERFLOW;
  }
  return FC5_OK;
}

// candidates per persistent launch (device memory 64 B each)
constexpr size_t kTwinBatch = 4u << 20;

}  // namespace

void fc5::free_twin(fc5_ctx* ctx) {
  for (void* p : {(void*)ctx->d_tri, (void*)ctx->d_twin_ctr, (void*)ctx->d_twin})
    if (p) (void)hipFree(p);
  if (ctx->h_twin_ctr) (void)hipHostFree(ctx->h_twin_ctr);
  ctx->d_tri = nullptr;
  ctx->d_twin_ctr = nullptr;
  ctx->d_twin = nullptr;
  ctx->h_twin_ctr = nullptr;
  ctx->d_twin_cap = 0;
}

extern "C" int fc5_twin_cands(uint64_t B, double u_lo, double u_hi, fc5_twin_cand* out, size_t cap, size_t* n) try {
  if (!n || (!out && cap)) return FC5_E_ARG;
  std::vector<TwinCandRaw> v;
  twin_cands(B, u_lo, u_hi, v);
  *n = v.size();
  for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = abi_of(v[i]);
  return FC5_OK;
} catch (...) {
  return abi_exception(nullptr);
}

extern "C" int fc5_twin_run(fc5_ctx* ctx, const fc5_twin_cand* cands, size_t n, fc5_twin_result* res) try {
  if (!ctx || !res || (!cands && n)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no pair table loaded");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::memset(res, 0, sizeof *res);
  ctx->twin_hits.clear();
  ctx->twin_sols.clear();
  std::vector<TwinCand> all(n);
  for (size_t i = 0; i < n; ++i) {
    std::string why;
    if (!pack_twin(raw_of(cands[i]), all[i], why))
      FAIL(ctx, FC5_E_ARG, "candidate " + std::to_string(i) + ": " + why);
  }
  // launches of at most kTwinBatch candidates and kMaxItems work items each
  for (size_t s = 0; s < n;) {
    size_t e = s;
    u64 items = 0;
    while (e < n && e - s < kTwinBatch) {
      const u64 rows = (u64)all[e].nres * (all[e].fmax1 / 64 + 1);
      if (e > s && items + rows > kMaxItems) break;
      items += rows;
      ++e;
    }
    std::vector<TwinCand> part(all.begin() + (long)s, all.begin() + (long)e);
    std::vector<u32> uid(e - s);
    for (size_t i = s; i < e; ++i) uid[i - s] = (u32)i;
    const int rc = twin_launch(ctx, part, uid, res);
    if (rc != FC5_OK) return rc;
    s = e;
  }
  const int rc = twin_verify(ctx, all);
  if (rc != FC5_OK) return rc;
  return finish(ctx, res, t0);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_twin_run_range(fc5_ctx* ctx, uint64_t B_lo, uint64_t B_hi, double u_lo, double u_hi,
                                  fc5_twin_result* res) try {
  if (!ctx || !res || B_lo > B_hi) return FC5_E_ARG;
  if (B_hi > 2353973) FAIL(ctx, FC5_E_ARG, "B_max capped at 2353973 (B^6 < 2^127)");
  std::vector<fc5_twin_cand> cands;
  std::vector<TwinCandRaw> v;
  for (u64 B = std::max<u64>(B_lo, 11); B <= B_hi; ++B) {
    if (!(B & 1) || B % 3 == 0 || B % 7 == 0) continue;  // Branch A (solve_617_v1.cu:1903-1907)
    v.clear();
    twin_cands(B, u_lo, u_hi, v);
    for (const TwinCandRaw& r : v) cands.push_back(abi_of(r));
  }
  return fc5_twin_run(ctx, cands.data(), cands.size(), res);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_twin_last_hits(fc5_ctx* ctx, fc5_twin_hit* out, size_t cap, size_t* n) try {
  if (!ctx || !n) return FC5_E_ARG;
  *n = ctx->twin_hits.size();
  if (out) std::copy_n(ctx->twin_hits.begin(), std::min(cap, ctx->twin_hits.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_twin_last_solutions(fc5_ctx* ctx, fc5_twin_solution* out, size_t cap, size_t* n) try {
  if (!ctx || !n) return FC5_E_ARG;
  *n = ctx->twin_sols.size();
  if (out) std::copy_n(ctx->twin_sols.begin(), std::min(cap, ctx->twin_sols.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}
