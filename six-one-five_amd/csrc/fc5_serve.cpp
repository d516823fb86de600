// fc5_serve.cpp -- resident serving loop of include/fc5.h (fc5_serve_*).
//
// One server thread per context owns the context's stream.  Submitted bands
// are packed on the caller's thread and queued; the server admits them in
// arrival order when a counter slot is free and their unit arrays fit into the
// resident unit arenas (first-fit ranges of two device arrays sized from the
// memory budget).  Every scheduling round builds a segment list that gives
// each admitted band an equal share of the round's grid points (its next work
// items) and launches the persistent kernel once over all of them; the
// per-slot counters and the round's hit count are copied back behind it on the
// same stream.  Two rounds are in flight: round k+1 is planned and enqueued
// before round k's results are taken apart (hits read back on a copy stream
// while round k+1 runs, demultiplexed by counter slot, work estimates
// updated), so the device does not wait for the host between rounds.  A band
// whose last items ran gives back its slot and arena ranges at once and goes
// to a finalizer thread, which verifies its hits (exact check on a stream of
// its own) and publishes its result.  Bands that arrive while others run join
// at the next planned round and share the device with them (continuous
// batching).  Counters are sums over work items and hits are sorted, so a
// band's result does not depend on what it shared rounds with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fc5_ctx.hpp"

using namespace fc5;

namespace {

using Clock = std::chrono::steady_clock;

constexpr u32 kSlotStride = 8;  // counter slots per band: index slot * 8 + class

// first-fit ranges over [0, cap)
struct Arena {
  u64 cap = 0, used = 0;
  std::map<u64, u64> free_;  // offset -> length
  void init(u64 c) {
    cap = c;
    used = 0;
    free_.clear();
    if (c) free_[0] = c;
  }
  bool alloc(u64 n, u64& off) {
    off = 0;
    if (!n) return true;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      off = it->first;
      const u64 rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      used += n;
      return true;
    }
    return false;
  }
  void release(u64 off, u64 n) {
    if (!n) return;
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

enum BandState { kPending, kActive, kFinalizing, kDone };

struct SBand {
  u64 id = 0;
  BandState state = kPending;
  int rc = FC5_OK;
  size_t n_in = 0, n_skipped = 0;
  std::vector<u8> cls_of;    // class of each caller index (0 = skipped)
  std::vector<int> slot_of;  // caller index -> index in units5 / units3
  std::vector<Cls5Unit> units5;
  std::vector<Cls234Unit> units3;
  std::vector<DevUnit5> dev5;
  std::vector<DevUnit3> dev3;
  u32 items5 = 0, items3 = 0;
  u32 next5 = 0, next3 = 0;  // items given to planned rounds
  u32 done5 = 0, done3 = 0;  // items of rounds that have completed
  double cost5 = 1, cost3 = 1;  // estimated grid points (c3 values) per work item
  u64 off5 = 0, off3 = 0;
  int slot = -1;
  u64 recs_seen = 0;  // hit records the kernel produced for this band so far
  u64 dropped = 0;    // of them, records beyond a round's hit_cap
  Counters ctr[8]{};  // the band's counter slots after its last round
  std::vector<fc5_hit> hits;
  std::vector<fc5_solution> sols;
  fc5_result res{};
  fc5_result cls_res[6]{};
  double t_submit = 0, t_admit = 0, t_done = 0, t_release = 0;
  u64 rounds = 0;
  double kernel_ms = 0;
  bool unplanned() const { return next5 < items5 || next3 < items3; }
  bool finished_items() const { return done5 >= items5 && done3 >= items3; }
};

// one scheduling round's buffers; two alternate so that one round can be
// planned and enqueued while the other still runs
struct Round {
  Segment* d_segs = nullptr;
  Segment* h_segs = nullptr;  // pinned
  HitRec* d_hits = nullptr;
  u64* d_nhits = nullptr;
  HitRec* h_hits = nullptr;   // pinned
  u64* h_nhits = nullptr;     // pinned
  Counters* h_ctr = nullptr;  // pinned snapshot of every counter slot after the round
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
  std::vector<std::pair<SBand*, std::pair<u32, u32>>> took;  // band, items of each kind
};

}  // namespace

struct fc5_server {
  fc5_ctx* ctx = nullptr;
  fc5_serve_opts opt{};
  int rows = kRowsPerItem;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stopping = false;
  int fail_rc = FC5_OK;
  u64 next_id = 1;
  std::deque<u64> pending;
  std::map<u64, std::unique_ptr<SBand>> bands;
  std::vector<SBand*> slots;  // admitted band per counter slot
  Arena a5, a3;
  DevUnit5* d5 = nullptr;
  DevUnit3* d3 = nullptr;
  Counters* d_ctr = nullptr;
  HitRec* d_hits2 = nullptr;  // second round's hit buffer (the first is the context's)
  u64* d_nhits2 = nullptr;
  Round rr[2];
  u64 planned = 0, processed = 0;  // rounds enqueued / taken apart; round k uses rr[k % 2]
  u32 max_fmax2 = 0, max_lim = 0;
  hipStream_t side = nullptr;  // Stage-1 emission of submitted ranges
  std::mutex side_mu;
  hipStream_t copy = nullptr;     // hit read-back beside the next round
  hipStream_t vstream = nullptr;  // exact checks of finished bands (finalizer thread)
  std::thread fin_th;
  std::deque<SBand*> fin_q;
  std::condition_variable cv_fin;
  bool fin_stop = false;
  u64 submitted = 0, admitted = 0, done = 0, rounds = 0, device_bytes = 0;
  double busy_ms = 0;
  Clock::time_point t0 = Clock::now();
  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void free_server_buffers(fc5_server* S) {
  for (void* p : {(void*)S->d5, (void*)S->d3, (void*)S->d_ctr, (void*)S->d_hits2, (void*)S->d_nhits2})
    if (p) (void)hipFree(p);
  S->d5 = nullptr;
  S->d3 = nullptr;
  S->d_ctr = nullptr;
  S->d_hits2 = nullptr;
  S->d_nhits2 = nullptr;
  for (Round& R : S->rr) {
    if (R.d_segs) (void)hipFree(R.d_segs);
    for (void* p : {(void*)R.h_segs, (void*)R.h_hits, (void*)R.h_nhits, (void*)R.h_ctr})
      if (p) (void)hipHostFree(p);
    for (hipEvent_t e : {R.ev0, R.ev1, R.ev_done})
      if (e) (void)hipEventDestroy(e);
    R = Round{};
  }
  for (hipStream_t* st : {&S->side, &S->copy, &S->vstream})
    if (*st) {
      (void)hipStreamDestroy(*st);
      *st = nullptr;
    }
}

// the band's result from its counter slots (same reduction as fc5_band_run)
void collect(SBand& B, bool strict) {
  const Counters* c = B.ctr;
  std::memset(&B.res, 0, sizeof B.res);
  std::memset(B.cls_res, 0, sizeof B.cls_res);
  for (int k = 2; k <= 5; ++k) {
    const Counters& C = c[k];
    fc5_result& r = B.cls_res[k];
    r.calls = C.calls;
    r.neg = C.neg;
    r.zero = C.zero;
    r.probes = C.probes;
    r.gated = C.pos - C.probes;
    r.maybe = C.maybe;
    r.side_maybe = C.side_maybe;
    r.jobs = C.jobs;
    r.fail_t = C.failT;
    r.strict_out = strict ? C.probes - C.queries : 0;
    B.res.calls += r.calls;
    B.res.neg += r.neg;
    B.res.zero += r.zero;
    B.res.probes += r.probes;
    B.res.gated += r.gated;
    B.res.maybe += r.maybe;
    B.res.side_maybe += r.side_maybe;
    B.res.jobs += r.jobs;
    B.res.fail_t += r.fail_t;
    B.res.strict_out += r.strict_out;
  }
  for (size_t i = 0; i < B.n_in; ++i)
    if (B.cls_of[i]) B.cls_res[B.cls_of[i]].units++;
}

// counters, verification and solutions of a band whose items have all run.
// Runs on the server thread; the caller publishes the band (state kDone)
// under the lock afterwards, and nothing else touches an open band's results.
void finalize(fc5_server* S, SBand& B) {
  fc5_ctx* ctx = S->ctx;
  collect(B, ctx->strict_gate != 0);
  std::sort(B.hits.begin(), B.hits.end(), [](const fc5_hit& x, const fc5_hit& y) {
    if (x.unit != y.unit) return x.unit < y.unit;
    if (x.e != y.e) return x.e < y.e;
    if (x.d != y.d) return x.d < y.d;
    return x.flags < y.flags;
  });
  if (ctx->verify && !B.hits.empty()) {
    const int N = (int)ctx->main.hdr.N;
    std::vector<VerifyJob> jobs;
    std::vector<size_t> hit_of;
    for (size_t i = 0; i < B.hits.size(); ++i) {
      const fc5_hit& h = B.hits[i];
      if (!(h.flags & 1u) || h.unit >= B.n_in) continue;
      const int k = B.cls_of[h.unit], sl = B.slot_of[h.unit];
      if (k == 5) {
        const Cls5Unit& U = B.units5[(size_t)sl];
        jobs.push_back(VerifyJob{5, U.B, U.u, U.lim, h.d, h.e, h.fp});
      } else if (k >= 2) {
        const Cls234Unit& U = B.units3[(size_t)sl];
        jobs.push_back(VerifyJob{k, U.B, U.u, U.lim, h.d, h.e, h.fp});
      } else {
        continue;
      }
      hit_of.push_back(i);
    }
    std::vector<PairTarget> targets;
    verify_targets(N, jobs, targets);
    std::vector<std::vector<std::pair<u32, u32>>> pairs;
    std::vector<u8> extra;
    std::vector<Solution> sols;
    const int prc = device_pairs(ctx, targets, pairs, &extra, S->vstream);
    if (prc == FC5_OK) {
      verify_finish_all(jobs, pairs, false, sols);
      for (size_t i = 0; i < extra.size(); ++i)
        if (extra[i]) B.hits[hit_of[i]].flags |= FC5_HIT_EXTRA;
    } else {
      B.rc = prc;
    }
    for (const Solution& x : sols) {
      fc5_solution o{};
      o.B = x.B;
      o.u = x.u;
      o.cls = x.cls;
      o.d = x.d;
      o.e = x.e;
      o.a = x.a;
      o.b = x.b;
      o.c = x.c;
      for (int k = 0; k < 5; ++k) o.terms[k] = x.terms[k];
      o.gcd = x.gcd;
      B.sols.push_back(o);
    }
  }
  B.res.n_hits = B.hits.size();
  B.res.hits_dropped = B.dropped;
  B.res.solutions = B.sols.size();
  B.res.units = B.units5.size() + B.units3.size();
  B.res.units_skipped = B.n_skipped;
  B.res.kernel_ms = B.kernel_ms;
  B.t_done = S->now();
  B.res.wall_ms = 1000.0 * (B.t_done - B.t_submit);
  if (B.dropped) B.rc = FC5_E_OVERFLOW;
  // the unit arrays are no longer needed
  B.dev5 = {};
  B.dev3 = {};
}

// admission in arrival order (under the lock); device uploads happen here,
// between rounds, on the server thread
int admit_locked(fc5_server* S) {
  fc5_ctx* ctx = S->ctx;
  while (!S->pending.empty()) {
    SBand& B = *S->bands.at(S->pending.front());
    if (B.items5 + B.items3 == 0) {  // nothing to run
      S->pending.pop_front();
      B.t_admit = B.t_release = S->now();
      finalize(S, B);
      B.state = kDone;
      S->done++;
      S->cv_done.notify_all();
      continue;
    }
    int slot = -1;
    for (size_t i = 0; i < S->slots.size(); ++i)
      if (!S->slots[i]) { slot = (int)i; break; }
    if (slot < 0) return FC5_OK;
    u64 off5 = 0, off3 = 0;
    if (!S->a5.alloc(B.dev5.size(), off5)) return FC5_OK;
    if (!S->a3.alloc(B.dev3.size(), off3)) {
      S->a5.release(off5, B.dev5.size());
      return FC5_OK;
    }
    S->pending.pop_front();
    B.slot = slot;
    B.off5 = off5;
    B.off3 = off3;
    for (DevUnit5& D : B.dev5) D.band = (u32)slot * kSlotStride + 5;
    for (DevUnit3& D : B.dev3) D.band = (u32)slot * kSlotStride + (D.band % kSlotStride);
    if (!B.dev5.empty())
      HIPCHK(ctx, hipMemcpy(S->d5 + off5, B.dev5.data(), B.dev5.size() * sizeof(DevUnit5), hipMemcpyHostToDevice));
    if (!B.dev3.empty())
      HIPCHK(ctx, hipMemcpy(S->d3 + off3, B.dev3.data(), B.dev3.size() * sizeof(DevUnit3), hipMemcpyHostToDevice));
    if (launch_fill_u64(reinterpret_cast<u64*>(S->d_ctr + (size_t)slot * kSlotStride),
                        kSlotStride * sizeof(Counters) / 8, 0, ctx->stream) != 0)
      FAIL(ctx, FC5_E_HIP, "counter reset");
    S->slots[(size_t)slot] = &B;
    B.state = kActive;
    B.t_admit = S->now();
    S->admitted++;
  }
  return FC5_OK;
}

// Plan one scheduling round over the bands with unplanned items and enqueue
// it on the context's stream into round buffers R: segment upload, hit-count
// and work-counter reset, the persistent kernel, then the counter snapshot and
// the hit count copied back; R.ev_done marks the end.
int plan_round(fc5_server* S, const std::vector<SBand*>& bands, Round& R) {
  fc5_ctx* ctx = S->ctx;
  const double share = S->opt.round_points / (double)std::max<size_t>(1, bands.size());
  u32 n_segs = 0, total = 0;
  R.took.clear();
  for (SBand* B : bands) {
    const u32 r5 = B->items5 - B->next5, r3 = B->items3 - B->next3;
    u32 n5 = 0, n3 = 0;
    double left = share;
    if (r5) {
      n5 = (u32)std::min<double>(r5, std::max(1.0, left / B->cost5));
      left -= n5 * B->cost5;
    }
    if (r3 && (left > 0 || !n5)) n3 = (u32)std::min<double>(r3, std::max(1.0, left / B->cost3));
    // one launch's items come from a 32-bit counter (kMaxItems per launch)
    n5 = (u32)std::min<u64>(n5, kMaxItems - total);
    n3 = (u32)std::min<u64>(n3, kMaxItems - total - n5);
    if (n5) {
      R.h_segs[n_segs++] = Segment{total, B->next5, (u32)B->off5, (u32)B->dev5.size(), 0, 0};
      total += n5;
    }
    if (n3) {
      R.h_segs[n_segs++] = Segment{total, B->next3, (u32)B->off3, (u32)B->dev3.size(), 1, 0};
      total += n3;
    }
    B->next5 += n5;
    B->next3 += n3;
    R.took.push_back({B, {n5, n3}});
  }
  BatchedArgs a{};
  a.units = S->d5;
  a.units3 = S->d3;
  a.segs = R.d_segs;
  a.n_segs = n_segs;
  a.total = total;
  a.gate = ctx->use_gate ? ctx->d_gate : ctx->d_gate_ones;
  a.ndg = ctx->d_ndg;
  a.ndg_K = ctx->ndg_K;
  a.use_gate = (u32)ctx->use_gate;
  a.strict = (u32)ctx->strict_gate;
  a.sieve = (u32)ctx->sweep;
  a.main = ctx->main.dev;
  a.side = ctx->side.present ? ctx->side.dev : XorDev{};
  a.hits = R.d_hits;
  a.hit_cap = ctx->hit_cap;
  a.nhits = R.d_nhits;
  a.ctr = S->d_ctr;
  a.item_ctr = ctx->d_item_ctr;  // rounds run in stream order, so one work counter serves both
  a.inv729 = inv_mod_2_64(729);
  const int bpc = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : batched_blocks_per_cu(S->rows, ctx->strict_gate != 0, ctx->sweep != 0);
  const int grid = (int)std::max<u64>(1, std::min<u64>((u64)ctx->num_cu * (u64)bpc, total));
  const size_t nslots = S->slots.size() * kSlotStride;
  HIPCHK(ctx, hipMemcpyAsync(R.d_segs, R.h_segs, n_segs * sizeof(Segment), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipEventRecord(R.ev0, ctx->stream));
  if (launch_fill_u64(R.d_nhits, 1, 0, ctx->stream) != 0 ||
      launch_fill_u64(reinterpret_cast<u64*>(ctx->d_item_ctr), 1, 0, ctx->stream) != 0)
    FAIL(ctx, FC5_E_HIP, "counter reset");
  if (launch_cls5_batched(a, grid, S->rows, ctx->stream) != 0) FAIL(ctx, FC5_E_HIP, "serving launch");
  HIPCHK(ctx, hipEventRecord(R.ev1, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(R.h_ctr, S->d_ctr, nslots * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(R.h_nhits, R.d_nhits, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipEventRecord(R.ev_done, ctx->stream));
  return FC5_OK;
}

// Take apart a completed round: read its hits back (on the copy stream, while
// the next round runs), hand them to their bands by counter slot, charge
// each band the records that did not fit the round's hit buffer, update the
// work estimates, and release the bands whose last items ran (their counters
// are kept, their slot and arena ranges are freed) to the finalizer.
int process_round(fc5_server* S, Round& R) {
  fc5_ctx* ctx = S->ctx;
  HIPCHK(ctx, hipEventSynchronize(R.ev_done));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, R.ev0, R.ev1));
  const u64 nh = *R.h_nhits, stored = std::min<u64>(nh, ctx->hit_cap);
  if (stored) {
    HIPCHK(ctx, hipMemcpyAsync(R.h_hits, R.d_hits, stored * sizeof(HitRec), hipMemcpyDeviceToHost, S->copy));
    HIPCHK(ctx, hipStreamSynchronize(S->copy));
  }
  std::vector<u64> kept(S->slots.size(), 0);
  for (u64 i = 0; i < stored; ++i) {
    const HitRec& h = R.h_hits[i];
    const u32 slot = (h.flags >> 8) / kSlotStride;
    if (slot < S->slots.size() && S->slots[slot]) {
      S->slots[slot]->hits.push_back(fc5_hit{h.unit, h.d, h.e, h.flags & 0xffu, h.fp});
      kept[slot]++;
    }
  }
  std::vector<SBand*> fin;
  for (auto& [B, n] : R.took) {
    B->done5 += n.first;
    B->done3 += n.second;
    const Counters* c = R.h_ctr + (size_t)B->slot * kSlotStride;
    u64 recs = 0;
    for (int k = 2; k <= 5; ++k) recs += c[k].recs;
    const u64 produced = recs - B->recs_seen, got = kept[(size_t)B->slot];
    B->dropped += produced > got ? produced - got : 0;  // the round's records of this band it did not keep
    B->recs_seen = recs;
    // running estimates of the work per item from the band's own counters
    if (B->done5) B->cost5 = std::max(1.0, (double)c[5].calls / B->done5);
    const double c3 = (double)(c[2].calls + c[3].calls + c[4].calls);
    if (B->done3 && c3 > 0) B->cost3 = std::max(1.0, c3 / B->done3);
    if (B->finished_items()) {
      std::memcpy(B->ctr, c, sizeof B->ctr);
      fin.push_back(B);
    }
  }
  std::lock_guard<std::mutex> lk(S->mu);  // what pollers and the stats read
  S->busy_ms += ms;
  S->rounds++;
  for (auto& [B, n] : R.took) {
    B->rounds++;
    B->kernel_ms += ms;
  }
  R.took.clear();
  for (SBand* B : fin) {
    B->state = kFinalizing;
    B->t_release = S->now();
    S->a5.release(B->off5, B->units5.size());
    S->a3.release(B->off3, B->units3.size());
    S->slots[(size_t)B->slot] = nullptr;
    S->fin_q.push_back(B);
  }
  if (!fin.empty()) S->cv_fin.notify_all();
  return FC5_OK;
}

void fail_open_bands_locked(fc5_server* S) {
  for (auto& [id, B] : S->bands)
    if (B->state == kPending || B->state == kActive) {
      B->rc = S->fail_rc;
      B->state = kDone;
    }
  S->pending.clear();
  S->cv_done.notify_all();
}

void serve_loop(fc5_server* S) {
  fc5_ctx* ctx = S->ctx;
  (void)hipSetDevice(ctx->device);
  for (;;) {
    std::vector<SBand*> plan;
    const bool inflight = S->planned > S->processed;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->fail_rc == FC5_OK) {
        const int rc = admit_locked(S);
        if (rc != FC5_OK) S->fail_rc = rc;
      }
      if (S->fail_rc != FC5_OK) {  // a device error ends the server: fail every open band
        fail_open_bands_locked(S);
        return;
      }
      for (SBand* B : S->slots)
        if (B && B->unplanned()) plan.push_back(B);
      if (plan.empty() && !inflight) {
        bool any = false;
        for (SBand* B : S->slots) any |= B != nullptr;
        if (S->stopping && S->pending.empty() && !any) return;
        S->cv_work.wait(lk);
        continue;
      }
    }
    int rc = FC5_OK;
    if (!plan.empty() && S->planned - S->processed < 2) {
      rc = plan_round(S, plan, S->rr[S->planned % 2]);
      if (rc == FC5_OK) S->planned++;
    } else {
      rc = process_round(S, S->rr[S->processed % 2]);
      if (rc == FC5_OK) S->processed++;
    }
    if (rc != FC5_OK) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->fail_rc = rc;
    }
  }
}

// finalizer thread: exact checks and results of the bands whose items have all run
void finalize_loop(fc5_server* S) {
  (void)hipSetDevice(S->ctx->device);
  for (;;) {
    SBand* B = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_fin.wait(lk, [&] { return !S->fin_q.empty() || S->fin_stop; });
      if (S->fin_q.empty()) return;
      B = S->fin_q.front();
      S->fin_q.pop_front();
    }
    finalize(S, *B);  // outside the lock: submitters and pollers are not held up
    std::lock_guard<std::mutex> lk(S->mu);
    B->state = kDone;
    S->done++;
    S->cv_done.notify_all();
  }
}

SBand* find_band(fc5_server* S, u64 id) {
  auto it = S->bands.find(id);
  return it == S->bands.end() ? nullptr : it->second.get();
}

}  // namespace

void fc5::serve_shutdown(fc5_ctx* ctx) {
  fc5_server* S = ctx->server;
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->fin_stop = true;
  }
  S->cv_fin.notify_all();
  if (S->fin_th.joinable()) S->fin_th.join();
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_server_buffers(S);
  delete S;
  ctx->server = nullptr;
}

extern "C" int fc5_serve_start(fc5_ctx* ctx, const fc5_serve_opts* opts) try {
  if (!ctx) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present || !ctx->main.hdr.N) FAIL(ctx, FC5_E_STATE, "serving needs a pair table with its N");
  if (ctx->verify && ctx->verify_table && (!ctx->d_exact || ctx->exact_N < (int)ctx->main.hdr.N))
    FAIL(ctx, FC5_E_STATE, "verify_table needs an exact table with N at least the pair table's N");
  // rounds hold a slice of each band: kRowsPerItem rows per item unless the context fixes it
  fc5_serve_opts o{16, ctx->rows_per_item > 0 ? ctx->rows_per_item : kRowsPerItem, 1ull << 23, 1ull << 22, 2e10,
                   288000000000ull};
  if (opts) {
    if (opts->max_bands) o.max_bands = opts->max_bands;
    if (opts->rows_per_item) o.rows_per_item = opts->rows_per_item;
    if (opts->arena_units5) o.arena_units5 = opts->arena_units5;
    if (opts->arena_units3) o.arena_units3 = opts->arena_units3;
    if (opts->round_points > 0) o.round_points = opts->round_points;
    if (opts->mem_budget) o.mem_budget = opts->mem_budget;
  }
  if (o.max_bands < 1 || o.max_bands > 64) FAIL(ctx, FC5_E_ARG, "max_bands must be 1..64");
  if (o.rows_per_item != 4 && o.rows_per_item != 8 && o.rows_per_item != 16 && o.rows_per_item != 32)
    FAIL(ctx, FC5_E_ARG, "rows_per_item must be 4, 8, 16 or 32");
  if (o.arena_units5 >= (1ull << 31) || o.arena_units3 >= (1ull << 31)) FAIL(ctx, FC5_E_ARG, "arena too large");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // every band the table can serve: B <= 42 (N + 1) gives fmax2 <= 2N + 2 and lim <= N + 1
  const u32 N = ctx->main.hdr.N;
  int rc = ensure_ndg(ctx, 2 * N + 2);
  if (rc == FC5_OK) rc = ensure_hits(ctx);
  if (rc != FC5_OK) return rc;
  const u64 table_bytes = ctx->main.hdr.m_cells * 8 + ctx->side.hdr.m_cells * 8 + ctx->n_exact * sizeof(u128);
  const u64 arena_bytes = o.arena_units5 * sizeof(DevUnit5) + o.arena_units3 * sizeof(DevUnit3);
  const u64 run_bytes = 2 * ctx->hit_cap * sizeof(HitRec) + (u64)o.max_bands * kSlotStride * sizeof(Counters);
  if (table_bytes + arena_bytes + run_bytes > o.mem_budget)
    FAIL(ctx, FC5_E_NOMEM, "table + unit arenas + run buffers exceed the memory budget");
  auto S = std::make_unique<fc5_server>();
  S->ctx = ctx;
  S->opt = o;
  S->rows = o.rows_per_item;
  S->slots.assign((size_t)o.max_bands, nullptr);
  S->a5.init(o.arena_units5);
  S->a3.init(o.arena_units3);
  S->max_fmax2 = 2 * N + 2;
  S->max_lim = N + 1;
  S->device_bytes = table_bytes + arena_bytes + run_bytes;
  const size_t nslots = (size_t)o.max_bands * kSlotStride;
  const size_t seg_bytes = 2 * (size_t)o.max_bands * sizeof(Segment);
  bool ok = hipMalloc(&S->d5, o.arena_units5 * sizeof(DevUnit5)) == hipSuccess &&
            hipMalloc(&S->d3, o.arena_units3 * sizeof(DevUnit3)) == hipSuccess &&
            hipMalloc(&S->d_ctr, nslots * sizeof(Counters)) == hipSuccess &&
            hipMalloc(&S->d_hits2, ctx->hit_cap * sizeof(HitRec)) == hipSuccess &&
            hipMalloc(&S->d_nhits2, 8) == hipSuccess;
  for (int k = 0; k < 2 && ok; ++k) {
    Round& R = S->rr[k];
    R.d_hits = k ? S->d_hits2 : ctx->d_hits;
    R.d_nhits = k ? S->d_nhits2 : ctx->d_nhits;
    ok = hipMalloc(&R.d_segs, seg_bytes) == hipSuccess &&
         hipHostMalloc((void**)&R.h_segs, seg_bytes, hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc((void**)&R.h_ctr, nslots * sizeof(Counters), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc((void**)&R.h_hits, ctx->hit_cap * sizeof(HitRec), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc((void**)&R.h_nhits, 8, hipHostMallocDefault) == hipSuccess &&
         hipEventCreate(&R.ev0) == hipSuccess && hipEventCreate(&R.ev1) == hipSuccess &&
         hipEventCreateWithFlags(&R.ev_done, hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    free_server_buffers(S.get());
    FAIL(ctx, FC5_E_NOMEM, "serving buffers");
  }
  if (launch_fill_u64(reinterpret_cast<u64*>(S->d_ctr), nslots * sizeof(Counters) / 8, 0, ctx->stream) != 0 ||
      hipStreamSynchronize(ctx->stream) != hipSuccess ||
      hipStreamCreateWithFlags(&S->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&S->vstream, hipStreamNonBlocking) != hipSuccess) {
    free_server_buffers(S.get());
    FAIL(ctx, FC5_E_HIP, "serving counter reset");
  }
  ctx->server = S.release();
  ctx->server->th = std::thread(serve_loop, ctx->server);
  ctx->server->fin_th = std::thread(finalize_loop, ctx->server);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_submit(fc5_ctx* ctx, const uint32_t* cls, const fc5_unit* units, size_t n,
                                uint64_t* band_id) try {
  if (!ctx || !band_id || (!units && n)) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  if (n > 0x7fffffffull) FAIL(ctx, FC5_E_ARG, "too many units in one band");
  auto B = std::make_unique<SBand>();
  B->t_submit = S->now();
  B->n_in = n;
  B->cls_of.assign(n, 0);
  B->slot_of.assign(n, -1);
  std::vector<Cls5Unit> P5;
  std::vector<Cls234Unit> P3;
  std::vector<char> ok5, ok3;
  std::vector<int> slot;
  pack_units(cls, reinterpret_cast<const u64*>(units), n, P5, ok5, P3, ok3, slot);
  for (size_t i = 0; i < n; ++i) {
    const int c = cls ? (int)cls[i] : 5;
    if (c == 5) {
      if (!ok5[(size_t)slot[i]]) { ++B->n_skipped; continue; }
      const Cls5Unit& U = P5[(size_t)slot[i]];
      if (!U.regular) FAIL(ctx, FC5_E_ARG, "unit " + std::to_string(i) + " is not a Stage-1 class-5 unit");
      if (U.fmax2 > S->max_fmax2) FAIL(ctx, FC5_E_ARG, "band needs a pair table with a larger N");
      B->cls_of[i] = 5;
      B->slot_of[i] = (int)B->units5.size();
      B->dev5.push_back(dev_unit5(U, (u32)i, 5));
      B->units5.push_back(U);
    } else if (c >= 2 && c <= 4) {
      if (!ok3[(size_t)slot[i]]) { ++B->n_skipped; continue; }
      const Cls234Unit& U = P3[(size_t)slot[i]];
      if (!U.regular) FAIL(ctx, FC5_E_ARG, "unit " + std::to_string(i) + " is not a Stage-1 class 2-4 unit");
      if (U.lim > S->max_lim) FAIL(ctx, FC5_E_ARG, "band needs a pair table with a larger N");
      B->cls_of[i] = (u8)c;
      B->slot_of[i] = (int)B->units3.size();
      B->dev3.push_back(dev_unit3(U, (u32)i, (u32)c));
      B->units3.push_back(U);
    } else {
      ++B->n_skipped;
    }
  }
  if (B->dev5.size() > S->a5.cap || B->dev3.size() > S->a3.cap)
    FAIL(ctx, FC5_E_NOMEM, "band larger than the resident unit arenas");
  if (!number_items(B->dev5, B->dev3, S->rows, B->items5, B->items3)) FAIL(ctx, FC5_E_ARG, kTooManyItems);
  u64 calls5 = 0, rows3 = 0, lim3 = 0;
  for (const Cls5Unit& U : B->units5) calls5 += cls5_calls(U);
  for (const DevUnit3& D : B->dev3) {
    rows3 += D.rows;
    lim3 += D.lim;
  }
  if (B->items5) B->cost5 = std::max(1.0, (double)calls5 / B->items5);
  // a d-row sweeps about (1 - 3^(-1/6)) * T^(1/6) ~ lim / 6 values of c3
  if (B->items3) B->cost3 = std::max(1.0, (double)rows3 / B->items3 * (double)lim3 / B->dev3.size() / 6.0);
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stopping) FAIL(ctx, FC5_E_STATE, "server is stopping");
  B->id = S->next_id++;
  *band_id = B->id;
  S->pending.push_back(B->id);
  S->bands[B->id] = std::move(B);
  S->submitted++;
  S->cv_work.notify_all();
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_submit_range(fc5_ctx* ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac,
                                      double hi_frac, uint64_t* band_id) try {
  if (!ctx || !band_id) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::vector<u32> cls;
  std::vector<fc5_unit> units;
  {
    std::lock_guard<std::mutex> lk(S->side_mu);
    const int rc = stage1_emit(ctx, S->side, lo, hi, cls_mask, lo_frac, hi_frac, cls, units);
    if (rc != FC5_OK) return rc;
  }
  return fc5_serve_submit(ctx, cls.data(), units.data(), units.size(), band_id);
} catch (...) {
  return abi_exception(ctx);
}

static int band_result(fc5_ctx* ctx, SBand* B, fc5_result* res) {
  if (B->state != kDone) return FC5_E_PENDING;
  if (res) *res = B->res;
  if (B->rc == FC5_E_OVERFLOW) ctx->err = "hit buffer overflow in a round of this band";
  else if (B->rc != FC5_OK) ctx->err = "the server failed while this band was open";
  return B->rc;
}

extern "C" int fc5_serve_poll(fc5_ctx* ctx, uint64_t band_id, fc5_result* res) try {
  if (!ctx) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  return band_result(ctx, B, res);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_wait(fc5_ctx* ctx, uint64_t band_id, double timeout_s, fc5_result* res) try {
  if (!ctx) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::unique_lock<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  auto ready = [&] { return B->state == kDone; };
  if (timeout_s < 0) S->cv_done.wait(lk, ready);
  else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
  return band_result(ctx, B, res);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_class_result(fc5_ctx* ctx, uint64_t band_id, int cls, fc5_result* out) try {
  if (!ctx || !out || cls < 2 || cls > 5) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  if (B->state != kDone) return FC5_E_PENDING;
  *out = B->cls_res[cls];
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_hits(fc5_ctx* ctx, uint64_t band_id, fc5_hit* out, size_t cap, size_t* n) try {
  if (!ctx || !n) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  if (B->state != kDone) return FC5_E_PENDING;
  *n = B->hits.size();
  if (out) std::copy_n(B->hits.begin(), std::min(cap, B->hits.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_solutions(fc5_ctx* ctx, uint64_t band_id, fc5_solution* out, size_t cap, size_t* n) try {
  if (!ctx || !n) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  if (B->state != kDone) return FC5_E_PENDING;
  *n = B->sols.size();
  if (out) std::copy_n(B->sols.begin(), std::min(cap, B->sols.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_timing_of(fc5_ctx* ctx, uint64_t band_id, fc5_serve_timing* out) try {
  if (!ctx || !out) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  *out = fc5_serve_timing{B->t_submit, B->t_admit, B->t_done, B->rounds, B->t_release};
  return B->state == kDone ? FC5_OK : FC5_E_PENDING;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_release(fc5_ctx* ctx, uint64_t band_id) try {
  if (!ctx) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  SBand* B = find_band(S, band_id);
  if (!B) FAIL(ctx, FC5_E_STATE, "unknown band");
  if (B->state != kDone) FAIL(ctx, FC5_E_PENDING, "band not finished");
  S->bands.erase(band_id);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_stats_get(fc5_ctx* ctx, fc5_serve_stats* out) try {
  if (!ctx || !out) return FC5_E_ARG;
  fc5_server* S = ctx->server;
  if (!S) FAIL(ctx, FC5_E_STATE, "no server running");
  std::lock_guard<std::mutex> lk(S->mu);
  u64 active = 0;
  for (SBand* B : S->slots) active += B != nullptr;
  *out = fc5_serve_stats{S->submitted, S->admitted, S->done, (u64)S->pending.size(), active, S->rounds,
                         S->a5.used, S->a3.used, S->device_bytes, S->busy_ms};
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_serve_stop(fc5_ctx* ctx) try {
  if (!ctx) return FC5_E_ARG;
  if (!ctx->server) FAIL(ctx, FC5_E_STATE, "no server running");
  const int rc = ctx->server->fail_rc;
  serve_shutdown(ctx);
  return rc;
} catch (...) {
  return abi_exception(ctx);
}
