// fc5_noalloc.cpp -- fc5_run_units_noalloc: the batched class-5 search of
// fc5_run_units for callers that cannot allocate device memory.
//
// Every device buffer the search needs beyond its inputs (the resident pair
// table and gate of the context) and its outputs (host arrays of the caller)
// lives in one file-scope device array, `scratch`, of kScratchBytes: the packed
// units, the launch segment, the d table, the counters, the work-queue and hit
// counters and the hit buffer.  Each GPU holds its own copy of the array until
// the process ends; its address on a device comes from hipGetSymbolAddress,
// called once per device when a context is created on it (noalloc_setup), and
// host code uses only that address, never the array's name.  The hits are
// checked exactly on the host (the reference's host verify,
// fourcore_cls5_gpu_v4.cu:513-542) with no device memory either.
//
// Calls on one GPU share its array (it keeps the last call's contents, so the
// counters are cleared by a fill kernel at the start of every call); a lock
// per device serializes them within the process, and calls from several
// processes on one GPU each use their own process's copy.  Set-up (the
// symbol's address, the host staging) happens once per device when a context
// is created on it; a call itself allocates no device memory and no pinned
// host memory (the unit packing uses ordinary host temporaries).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "fc5_ctx.hpp"

using namespace fc5;

namespace {

constexpr u32 kNaUnits = FC5_NOALLOC_UNITS;
constexpr u32 kNaHits = FC5_NOALLOC_HITS;
constexpr u32 kNaDmax = (1u << 18) - 1;  // largest free term d: fmax2 = (B - 1) / 21
constexpr u32 kNaK = kNaDmax / 64 + 1;   // d-table columns per residue class
constexpr int kMaxDevices = 64;

struct Scratch {
  Counters ctr[16];
  u64 nhits;
  u64 item_ctr;  // the kernel's 32-bit work counter, cleared as one word
  Segment seg;
  u32 pad[2];
  DevUnit5 units[kNaUnits];
  HitRec hits[kNaHits];
  u32 ndg[64 * kNaK];
};
constexpr size_t kScratchBytes = (sizeof(Scratch) + 255) / 256 * 256;
static_assert(kScratchBytes <= FC5_NOALLOC_BYTES, "scratch array larger than the stated maximum");

// host staging of the same sizes, guarded by the per-device lock
struct HostStage {
  DevUnit5 units[kNaUnits];
  HitRec raw[kNaHits];
  fc5_hit hits[kNaHits];
  u32 ndg[64 * kNaK];
  Counters ctr[16];
  u64 nhits;
};

struct DeviceScratch {
  std::mutex mu;
  Scratch* ptr = nullptr;  // device address of `scratch` on this device (hipGetSymbolAddress)
  int setup_rc = FC5_E_STATE;
  bool set_up = false;
  u32 ndg_K = 0;  // d-table columns the array holds (0: none)
  HostStage* stage = nullptr;
};

DeviceScratch g_dev[kMaxDevices];

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Called with `device` current by fc5_ctx_create, once per device per process.
int fc5::noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return FC5_E_ARG;
  DeviceScratch& D = g_dev[device];
  std::lock_guard<std::mutex> lk(D.mu);
  if (D.set_up) return D.setup_rc;
  D.set_up = true;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    D.setup_rc = FC5_E_HIP;
    return D.setup_rc;
  }
  D.ptr = static_cast<Scratch*>(p);
  D.stage = new HostStage;  // host staging of the entry point, kept until the process ends
  D.setup_rc = FC5_OK;
  return FC5_OK;
}

// every a <= b <= hi with a^6 + b^6 = T (two pointers over x^6), for each the
// class-5 identity and the solution record; no memory is allocated
static void verify_hit(u64 B, u64 u, u32 hi, u32 d, u32 e, u64 fp, fc5_solution* sols, size_t cap, size_t& n) {
  u128 T;
  if (!exact_T_cls5(B, u, d, e, T) || (u64)T != fp) return;
  auto p6 = [](u64 x) {
    const u128 x2 = (u128)x * x;
    return x2 * x2 * x2;
  };
  u64 b = std::min<u64>(hi, iroot6(T));
  for (u64 a = 1; a <= b; ++a) {
    const u128 pa = p6(a);
    if (pa > T) break;
    while (b >= a && pa + p6(b) > T) --b;
    if (b < a) break;
    if (pa + p6(b) != T || !verify_cls5(B, u, d, e, a, b)) continue;
    const i64 t[5] = {42LL * (i64)a, 42LL * (i64)b, 21LL * (i64)d, 14LL * (i64)e, (i64)u};
    Solution s;
    if (!solution_record(B, u, 5, d, e, (u32)a, (u32)b, 0, t, s)) continue;
    if (n < cap) {
      fc5_solution& o = sols[n];
      o = fc5_solution{};
      o.B = s.B;
      o.u = s.u;
      o.cls = s.cls;
      o.d = s.d;
      o.e = s.e;
      o.a = s.a;
      o.b = s.b;
      o.c = s.c;
      for (int k = 0; k < 5; ++k) o.terms[k] = s.terms[k];
      o.gcd = s.gcd;
    }
    ++n;
  }
}

extern "C" int fc5_run_units_noalloc(fc5_ctx* ctx, const fc5_unit* units, size_t n, fc5_result* res, fc5_hit* hits,
                                     size_t hit_cap, size_t* n_hits, fc5_solution* sols, size_t sol_cap,
                                     size_t* n_sols) try {
  if (!ctx || !res || !n_hits || !n_sols || (!units && n) || (!hits && hit_cap) || (!sols && sol_cap))
    return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no pair table loaded");
  if (ctx->device < 0 || ctx->device >= kMaxDevices) FAIL(ctx, FC5_E_ARG, "device index out of range");
  if (n > kNaUnits) FAIL(ctx, FC5_E_ARG, "more than FC5_NOALLOC_UNITS units: split the call or use fc5_run_units");
  DeviceScratch& D = g_dev[ctx->device];
  std::lock_guard<std::mutex> lk(D.mu);
  if (!D.set_up || D.setup_rc != FC5_OK || !D.ptr || !D.stage)
    FAIL(ctx, FC5_E_HIP, "the scratch array's device address is not available (hipGetSymbolAddress failed)");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HostStage& H = *D.stage;
  Scratch* S = D.ptr;
  std::memset(res, 0, sizeof *res);
  // pack (the reference's pack_unit5, fourcore_cls5_gpu_v4.cu:169-184) into T-space units
  u32 n_dev = 0, dmax = 0;
  u64 skipped = 0;
  for (size_t i = 0; i < n; ++i) {
    Cls5Unit U;
    if (!pack_cls5(units[i].B, units[i].u, U)) {
      ++skipped;
      continue;
    }
    if (!U.regular) FAIL(ctx, FC5_E_ARG, "unit " + std::to_string(i) + " is not a Stage-1 class-5 unit");
    if (U.fmax2 > kNaDmax) FAIL(ctx, FC5_E_ARG, "B above FC5_NOALLOC_BMAX");
    dmax = std::max(dmax, U.fmax2);
    H.units[n_dev++] = dev_unit5(U, (u32)i, 5);
  }
  u64 rows = 0;
  for (u32 k = 0; k < n_dev; ++k) rows += H.units[k].rows;
  const int E = ctx->rows_per_item > 0 ? ctx->rows_per_item : auto_rows_per_item(ctx->num_cu, rows, ctx->strict_gate != 0, ctx->sweep != 0);
  u64 items = 0;
  for (u32 k = 0; k < n_dev; ++k) {
    H.units[k].item_base = (u32)items;
    items += (H.units[k].rows + (u32)E - 1) / (u32)E;
  }
  if (items > kMaxItems) FAIL(ctx, FC5_E_ARG, kTooManyItems);
  // the d table (the context's ensure_ndg, into the array), uploaded when it must grow
  const u32 K = dmax / 64 + 1;
  if (n_dev && D.ndg_K < K) {
    for (u32 r = 0; r < 64; ++r)
      for (u32 k = 0; k < K; ++k) {
        const u64 dd = (u64)r + 64ull * k;
        const u64 d3 = dd * dd * dd;
        const u128 dq = ((u128)d3 * d3) >> 6;
        H.ndg[(size_t)r * K + k] = (u32)(kGateModU - (u64)(dq % kGateModU));
      }
    D.ndg_K = 0;  // until the copy below has completed
    HIPCHK(ctx, hipMemcpyAsync(S->ndg, H.ndg, (size_t)64 * K * 4, hipMemcpyHostToDevice, ctx->stream));
  }
  const u32 ndgK = n_dev ? std::max(K, D.ndg_K) : 0;
  const Segment seg{0, 0, 0, n_dev, 0, 0};
  if (n_dev) {
    HIPCHK(ctx, hipMemcpyAsync(S->units, H.units, (size_t)n_dev * sizeof(DevUnit5), hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(&S->seg, &seg, sizeof seg, hipMemcpyHostToDevice, ctx->stream));
  }
  // the array keeps the last call's counters: cleared by a fill kernel (valid under capture too)
  if (launch_fill_u64(reinterpret_cast<u64*>(S->ctr), sizeof S->ctr / 8, 0, ctx->stream) != 0 ||
      launch_fill_u64(&S->nhits, 1, 0, ctx->stream) != 0 || launch_fill_u64(&S->item_ctr, 1, 0, ctx->stream) != 0)
    FAIL(ctx, FC5_E_HIP, "counter reset");
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (items) {
    BatchedArgs a{};
    a.units = S->units;
    a.segs = &S->seg;
    a.n_segs = 1;
    a.total = (u32)items;
    a.gate = ctx->use_gate ? ctx->d_gate : ctx->d_gate_ones;
    a.ndg = S->ndg;
    a.ndg_K = ndgK;
    a.use_gate = (u32)ctx->use_gate;
    a.strict = (u32)ctx->strict_gate;
    a.sieve = (u32)ctx->sweep;
    a.main = ctx->main.dev;
    a.side = ctx->side.present ? ctx->side.dev : XorDev{};
    a.hits = S->hits;
    a.hit_cap = kNaHits;
    a.nhits = &S->nhits;
    a.ctr = S->ctr;
    a.item_ctr = reinterpret_cast<u32*>(&S->item_ctr);
    a.inv729 = inv_mod_2_64(729);
    a.units3 = nullptr;
    const int bpc = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : batched_blocks_per_cu(E, ctx->strict_gate != 0, ctx->sweep != 0);
    const int grid = (int)std::max<u64>(1, std::min<u64>((u64)ctx->num_cu * (u64)bpc, items));
    if (launch_cls5_batched(a, grid, E, ctx->stream) != 0) FAIL(ctx, FC5_E_HIP, "batched launch");
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(H.ctr, S->ctr, sizeof H.ctr, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&H.nhits, &S->nhits, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (n_dev && K > D.ndg_K) D.ndg_K = K;
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const u64 nh = H.nhits, stored = std::min<u64>(nh, kNaHits);
  if (stored) HIPCHK(ctx, hipMemcpy(H.raw, S->hits, stored * sizeof(HitRec), hipMemcpyDeviceToHost));
  // counters as fc5_band_run reduces them (batched kernel: gated = positives - probes)
  const Counters& C = H.ctr[5];
  res->calls = C.calls;
  res->neg = C.neg;
  res->zero = C.zero;
  res->probes = C.probes;
  res->gated = C.pos - C.probes;
  res->maybe = C.maybe;
  res->side_maybe = C.side_maybe;
  res->strict_out = ctx->strict_gate ? C.probes - C.queries : 0;
  res->units = n_dev;
  res->units_skipped = skipped;
  res->kernel_ms = ms;
  for (u64 i = 0; i < stored; ++i) {
    const HitRec& h = H.raw[i];
    H.hits[i] = fc5_hit{h.unit, h.d, h.e, h.flags & 0xffu, h.fp};
  }
  std::sort(H.hits, H.hits + stored, [](const fc5_hit& x, const fc5_hit& y) {
    if (x.unit != y.unit) return x.unit < y.unit;
    if (x.e != y.e) return x.e < y.e;
    if (x.d != y.d) return x.d < y.d;
    return x.flags < y.flags;
  });
  std::copy_n(H.hits, std::min<u64>(stored, hit_cap), hits);
  *n_hits = (size_t)stored;
  res->n_hits = stored;
  res->hits_dropped = nh - stored;
  size_t ns = 0;
  if (ctx->verify) {
    const u32 N = ctx->main.hdr.N;
    for (u64 i = 0; i < stored; ++i) {
      const fc5_hit& h = H.hits[i];
      if (!(h.flags & 1u) || h.unit >= n) continue;
      const u64 B = units[h.unit].B, u = units[h.unit].u;
      const u32 lim = (u32)((B - 1) / 42);
      verify_hit(B, u, N ? std::min(lim, N) : lim, h.d, h.e, h.fp, sols, sol_cap, ns);
    }
  }
  *n_sols = ns;
  res->solutions = ns;
  res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (nh > stored) FAIL(ctx, FC5_E_OVERFLOW, "more than FC5_NOALLOC_HITS hits: the rest were dropped");
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}
