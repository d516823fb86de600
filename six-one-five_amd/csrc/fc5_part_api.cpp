// fc5_part_api.cpp -- partitioned xor-filter probing (fc5_part.hip) in the
// batched search, and its measurement against direct queries of the same keys.
//
// A batched run with partitioned probing:
//   1. emit launches of the persistent kernel (probe mode 1): the gate
//      survivors' hashes go into 4096-hash chunks of one flat buffer; a launch
//      takes no new work item once 85% of the chunks are taken, and a probe
//      that finds no free chunk is queried at once;
//   2. after each emit launch the partition pipeline over the filled chunks:
//      the accepted hashes (exact: every candidate gets the full query) are
//      appended to one list; then the next emit launch continues with the
//      work items left;
//   3. one identify launch of the persistent kernel (probe mode 2) over the
//      whole band: every probe again, the main table's answer being membership
//      of its hash in the sorted accepted set.  Its counters and hit records
//      are the run's, the same as with a gather per cell of every probe.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "fc5_ctx.hpp"
#include "fc5_part.hpp"

namespace {

using namespace fc5;

// device allocations freed together at scope exit
struct DevArena {
  std::vector<void*> p;
  template <class T>
  T* get(u64 count) {
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T*)q;
  }
  ~DevArena() {
    for (void* q : p) (void)hipFree(q);
  }
};

// records per bucket for an expected mean: 8 standard deviations and a tile's
// worth above it (a bucket's count is binomial with mean `mean`)
u64 bucket_cap(double mean) { return (u64)std::ceil(mean + 8.0 * std::sqrt(mean + 1.0)) + 4096; }

}  // namespace

namespace fc5 {

struct PartBufs {
  u64 records = 0;  // emit capacity in hashes: n_chunks * kEmitChunk
  u32 n_chunks = 0;
  PartGeom g{};
  low_t* lo = nullptr;
  u64* flat = nullptr;
  u32* chunk_fill = nullptr;
  u32* ctr = nullptr;  // [0] chunk counter, [1] side n, [2] side lost, [3] acc n, [4] acc lost
  u64 *ch = nullptr, *fh = nullptr, *side_h = nullptr, *acc_h = nullptr, *set_h = nullptr;
  low_t *ca = nullptr, *fa = nullptr;
  u32 *cfill = nullptr, *ffill = nullptr, *set_bits = nullptr;
  u64 cap_c = 0, cap_f = 0;
  u32 side_cap = 0, acc_cap = 0, set_cap = 0;
  hipEvent_t ev[16] = {};  // start and end of each of the pipeline's eight stages
  std::vector<void*> owned;
  ~PartBufs() {
    for (void* q : owned) (void)hipFree(q);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  template <class T>
  bool get(T*& p, u64 count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<u64>(count, 1) * sizeof(T)) != hipSuccess) return false;
    owned.push_back(q);
    p = (T*)q;
    return true;
  }
};

void free_part(fc5_ctx* ctx) {
  if (!ctx->part) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx->part;
  ctx->part = nullptr;
}

bool use_partitioned(const fc5_ctx* ctx) {
  if (ctx->strict_gate || !ctx->main.present || !ctx->main.cells) return false;
  if (ctx->probe_mode >= 0) return ctx->probe_mode == 1;
  return ctx->main.hdr.m_cells >= (1ull << 28) && ctx->main.hdr.r >= 24;
}

namespace {

// the buffers for the current main table (made once per table and capacity)
int ensure_part(fc5_ctx* ctx, PartBufs*& out) {
  if (ctx->part) {
    out = ctx->part;
    return FC5_OK;
  }
  const XorDev& x = ctx->main.dev;
  PartBufs* P = new PartBufs;
  P->n_chunks = (u32)std::max<u64>(1, ctx->part_records / kEmitChunk);
  P->records = (u64)P->n_chunks * kEmitChunk;
  P->g = make_part_geom(x, nullptr);
  if (P->g.nc == 0 || P->g.nc > kPartMaxBins) {
    delete P;
    FAIL(ctx, FC5_E_ARG, "table too large for partitioned probing (two partition levels)");
  }
  const u32 nf = P->g.nc << kFineShift;
  const u64 ncr = (u64)P->g.nc * kCoarseSubs;
  const double per_slice = (double)P->records * (double)kSliceCells / (double)x.block;
  P->cap_c = bucket_cap(per_slice * (1u << kFineShift) / kCoarseSubs);
  P->cap_f = bucket_cap(per_slice);
  // candidates: about 2^-min(r, kLowBits) of the records, so the list holds twice that
  const u32 lowb = std::min<u32>(ctx->main.hdr.r, kLowBits);
  P->side_cap = (u32)std::min<u64>(0x7FFFFFFFull, (P->records >> lowb) * 2 + (1u << 20));
  P->acc_cap = 1u << 22;
  const u64 need = 3 * P->g.lo_third * sizeof(low_t) + P->records * 8 + (u64)P->n_chunks * 4 + 64 +
                   (P->cap_c * ncr + P->cap_f * nf) * (8 + sizeof(low_t)) + (ncr * kCoarseFillStride + nf) * 4 +
                   ((u64)P->side_cap + P->acc_cap) * 8 + (1u << kSetFilterBits) / 8;
  if (ctx->part_hbm_limit && need > ctx->part_hbm_limit) {
    delete P;
    FAIL(ctx, FC5_E_NOMEM, "partitioned probing buffers exceed the option part_hbm_limit");
  }
  const bool ok = P->get(P->lo, 3 * P->g.lo_third) && P->get(P->flat, P->records) &&
                  P->get(P->chunk_fill, P->n_chunks) && P->get(P->ctr, 8) && P->get(P->ch, P->cap_c * ncr) &&
                  P->get(P->ca, P->cap_c * ncr) && P->get(P->cfill, ncr * kCoarseFillStride) &&
                  P->get(P->fh, P->cap_f * nf) && P->get(P->fa, P->cap_f * nf) && P->get(P->ffill, nf) &&
                  P->get(P->side_h, P->side_cap) && P->get(P->acc_h, P->acc_cap) && P->get(P->set_bits, (1u << kSetFilterBits) / 32);
  if (!ok) {
    delete P;
    FAIL(ctx, FC5_E_NOMEM, "partitioned probing buffers (lower the option part_records)");
  }
  P->g.lo = P->lo;
  for (hipEvent_t& e : P->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete P;
      FAIL(ctx, FC5_E_HIP, "pipeline events");
    }
  if (launch_part_low(P->g, P->lo, ctx->stream) != 0) {
    delete P;
    FAIL(ctx, FC5_E_HIP, "low-bits table");
  }
  ctx->part = out = P;
  return FC5_OK;
}

// the partition pipeline over the filled emit chunks: accepted hashes appended
// to P->acc_h (P->ctr[3])
int run_pipeline(fc5_ctx* ctx, PartBufs* P) {
  hipStream_t s = ctx->stream;
  const PartGeom& g = P->g;
  const u32 nf = g.nc << kFineShift;
  const u64 ncr = (u64)g.nc * kCoarseSubs;
  const PartSet coarse0{P->ch, nullptr, P->cfill, P->cap_c, g.nc, kCoarseSubs, kCoarseFillStride};
  const PartSet coarse{P->ch, P->ca, P->cfill, P->cap_c, g.nc, kCoarseSubs, kCoarseFillStride};
  const PartSet fine0{P->fh, nullptr, P->ffill, P->cap_f, nf, 1, 1};
  const PartSet fine{P->fh, P->fa, P->ffill, P->cap_f, nf, 1, 1};
  const PartSide side{P->side_h, P->ctr + 1, P->ctr + 2, P->side_cap};
  const PartSide acc{P->acc_h, P->ctr + 3, P->ctr + 4, P->acc_cap};
  // stage k (first, fine 0, lookup 0, fine 1, lookup 1, fine 2, last, check)
  // between events 2k and 2k + 1; the counters a stage appends to are
  // cleared before its start event, so the events bracket the kernel alone
  auto stage = [&](int k, bool launched) {
    return launched && hipEventRecord(P->ev[2 * k + 1], s) == hipSuccess;
  };
  auto start = [&](int k) { return hipEventRecord(P->ev[2 * k], s) == hipSuccess; };
  bool ok = hipMemsetAsync(P->ctr + 1, 0, 4, s) == hipSuccess &&
            hipMemsetAsync(P->cfill, 0, ncr * kCoarseFillStride * 4, s) == hipSuccess && start(0) &&
            stage(0, launch_part_first(P->flat, P->records, P->chunk_fill, g, coarse0, side, s) == 0);
  for (int p = 0; p < 3 && ok; ++p) {
    const PartSet& fout = p == 0 ? fine0 : fine;
    ok = hipMemsetAsync(P->ffill, 0, (u64)nf * 4, s) == hipSuccess && start(1 + 2 * p) &&
         stage(1 + 2 * p, launch_part_fine(p, g, p == 0 ? coarse0 : coarse, fout, side, s) == 0);
    if (ok && p < 2)
      ok = hipMemsetAsync(P->cfill, 0, ncr * kCoarseFillStride * 4, s) == hipSuccess && start(2 + 2 * p) &&
           stage(2 + 2 * p, launch_part_lookup(p, g, fout, coarse, side, s) == 0);
    else if (ok)
      ok = start(6) && stage(6, launch_part_last(g, fine, side, s) == 0);
  }
  ok = ok && start(7) && stage(7, launch_part_check(g, side, acc, ctx->num_cu * 4, s) == 0);
  if (!ok) FAIL(ctx, FC5_E_HIP, "partition pipeline launch");
  return FC5_OK;
}

// after the pipeline's synchronize: its stage times added to the run's
int add_stage_times(fc5_ctx* ctx, PartBufs* P) {
  for (int k = 0; k < 8; ++k) {
    float t = 0;
    HIPCHK(ctx, hipEventElapsedTime(&t, P->ev[2 * k], P->ev[2 * k + 1]));
    ctx->part_stage_ms[k] += t;
  }
  return FC5_OK;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int run_partitioned(fc5_ctx* ctx, const BatchedArgs& a0, int grid, int rows) {
  PartBufs* P = nullptr;
  int rc = ensure_part(ctx, P);
  if (rc == FC5_E_NOMEM && ctx->probe_mode < 0) {
    // chosen by table size only: without room for the buffers, probe with gathers
    (void)hipGetLastError();
    if (launch_cls5_batched(a0, grid, rows, ctx->stream) != 0) FAIL(ctx, FC5_E_HIP, "batched launch");
    return FC5_OK;
  }
  if (rc != FC5_OK) return rc;
  hipStream_t s = ctx->stream;
  BatchedArgs a = a0;
  a.pmode = 1;
  a.flat = P->flat;
  a.chunk_ctr = P->ctr;
  a.chunk_fill = P->chunk_fill;
  a.n_chunks = P->n_chunks;
  a.chunk_limit = (u32)std::max<u64>(1, (u64)P->n_chunks * 85 / 100);
  a.acc_h = P->acc_h;
  a.acc_n = P->ctr + 3;
  a.acc_lost = P->ctr + 4;
  a.acc_cap = P->acc_cap;
  std::fill(std::begin(ctx->part_ms), std::end(ctx->part_ms), 0.0);
  std::fill(std::begin(ctx->part_stage_ms), std::end(ctx->part_stage_ms), 0.0);
  std::fill(std::begin(ctx->part_counts), std::end(ctx->part_counts), 0);
  HIPCHK(ctx, hipMemsetAsync(P->ctr, 0, 8 * 4, s));
  u32 h_ctr[8];
  u64 item_ctr = 0;
  // the pieces: emit launch, then the partition pipeline over its chunks
  for (u64 piece = 0;; ++piece) {
    auto t0 = std::chrono::steady_clock::now();
    if (hipMemsetAsync(P->ctr, 0, 4, s) != hipSuccess ||
        hipMemsetAsync(P->chunk_fill, 0, (u64)P->n_chunks * 4, s) != hipSuccess ||
        launch_cls5_batched(a, grid, rows, s) != 0 || hipStreamSynchronize(s) != hipSuccess)
      FAIL(ctx, FC5_E_HIP, "emit launch");
    ctx->part_ms[0] += ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    rc = run_pipeline(ctx, P);
    if (rc != FC5_OK) return rc;
    HIPCHK(ctx, hipMemcpyAsync(h_ctr, P->ctr, sizeof h_ctr, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&item_ctr, a.item_ctr, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->part_ms[1] += ms_since(t0);
    rc = add_stage_times(ctx, P);
    if (rc != FC5_OK) return rc;
    ctx->part_counts[0] = piece + 1;
    ctx->part_counts[2] += std::min(h_ctr[1], P->side_cap);
    ctx->part_counts[3] += (u64)std::min(h_ctr[0], P->n_chunks);
    if (h_ctr[2] || h_ctr[4])
      FAIL(ctx, FC5_E_OVERFLOW, "partitioned probing: candidate or accepted list full (use probe_mode 0)");
    if ((u32)item_ctr >= a.total) break;
  }
  // the accepted set, sorted and unique, with its filter of top bits
  const u32 n_acc = h_ctr[3];
  std::vector<u64> acc(n_acc);
  if (n_acc) HIPCHK(ctx, hipMemcpy(acc.data(), P->acc_h, (u64)n_acc * 8, hipMemcpyDeviceToHost));
  std::sort(acc.begin(), acc.end());
  acc.erase(std::unique(acc.begin(), acc.end()), acc.end());
  std::vector<u32> bits((1u << kSetFilterBits) / 32, 0);
  for (u64 h : acc) {
    const u32 b = (u32)(h >> (64 - kSetFilterBits));
    bits[b >> 5] |= 1u << (b & 31u);
  }
  if (acc.size() > P->set_cap) {
    u64* nh = nullptr;
    if (!P->get(nh, acc.size())) FAIL(ctx, FC5_E_NOMEM, "accepted set");
    P->set_h = nh;
    P->set_cap = (u32)acc.size();
  }
  if (!acc.empty()) HIPCHK(ctx, hipMemcpy(P->set_h, acc.data(), acc.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(P->set_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
  ctx->part_counts[1] = acc.size();
  // the identify launch: counters and hits of the run
  auto t0 = std::chrono::steady_clock::now();
  BatchedArgs b = a0;
  b.pmode = 2;
  b.set_h = P->set_h;
  b.set_n = (u32)acc.size();
  b.set_bits = P->set_bits;
  if (launch_fill_u64((u64*)ctx->d_ctr, 16 * sizeof(Counters) / 8, 0, s) != 0 ||
      launch_fill_u64(ctx->d_nhits, 1, 0, s) != 0 || launch_fill_u64((u64*)ctx->d_item_ctr, 1, 0, s) != 0 ||
      launch_cls5_batched(b, grid, rows, s) != 0 || hipStreamSynchronize(s) != hipSuccess)
    FAIL(ctx, FC5_E_HIP, "identify launch");
  ctx->part_ms[2] = ms_since(t0);
  return FC5_OK;
}

}  // namespace fc5

extern "C" int fc5_part_stats(fc5_ctx* ctx, double* ms, uint64_t* counts) try {
  if (!ctx || !ms || !counts) return FC5_E_ARG;
  for (int i = 0; i < 3; ++i) ms[i] = ctx->part_ms[i];
  for (int i = 0; i < 8; ++i) ms[3 + i] = ctx->part_stage_ms[i];
  for (int i = 0; i < 4; ++i) counts[i] = ctx->part_counts[i];
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_measure_partitioned(fc5_ctx* ctx, uint64_t n, double* ms, uint64_t* counts) try {
  if (!ctx || !ms || !counts || n == 0 || n > 0xFFFFFFFFull) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present || !ctx->main.cells) FAIL(ctx, FC5_E_STATE, "no device table");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const XorDev& x = ctx->main.dev;
  DevArena A;
  const PartGeom g0 = make_part_geom(x, nullptr);
  low_t* lo = A.get<low_t>(3 * g0.lo_third);
  if (!lo) FAIL(ctx, FC5_E_NOMEM, "low-bits table");
  const PartGeom g = make_part_geom(x, lo);
  const u64 ns = (x.block + kSliceCells - 1) >> kSliceShift;
  if (g.nc == 0 || g.nc > kPartMaxBins) FAIL(ctx, FC5_E_ARG, "table too large for two partition levels");
  const u32 nf = g.nc << kFineShift;
  // expected records per coarse bucket and per slice (uniform cell positions)
  const double per_slice = (double)n * (double)kSliceCells / (double)x.block;
  const u64 cap_c = bucket_cap(per_slice * (1u << kFineShift) / kCoarseSubs);  // per region
  const u64 cap_f = bucket_cap(per_slice);
  const u32 side_cap = (u32)std::min<u64>(
      0x7FFFFFFFull, (n >> std::min<u32>(ctx->main.hdr.r, kLowBits)) * 2 + (1u << 20));
  u64* flat = A.get<u64>(n);
  const u64 ncr = (u64)g.nc * kCoarseSubs;  // coarse regions
  u64* ch = A.get<u64>(cap_c * ncr);
  low_t* ca = A.get<low_t>(cap_c * ncr);
  u32* cfill = A.get<u32>(ncr * kCoarseFillStride);
  u64* fh = A.get<u64>(cap_f * nf);
  low_t* fa = A.get<low_t>(cap_f * nf);
  u32* ffill = A.get<u32>(nf);
  u64* sh = A.get<u64>(side_cap);
  u64* ah = A.get<u64>(side_cap);
  u32* ctr = A.get<u32>(4);  // side n, side lost, accepted n, accepted lost
  unsigned long long* sums = A.get<unsigned long long>(4);
  if (!flat || !ch || !ca || !cfill || !fh || !fa || !ffill || !sh || !ah || !ctr || !sums)
    FAIL(ctx, FC5_E_NOMEM, "partition buffers");
  const PartSet coarse0{ch, nullptr, cfill, cap_c, g.nc, kCoarseSubs, kCoarseFillStride};
  const PartSet coarse{ch, ca, cfill, cap_c, g.nc, kCoarseSubs, kCoarseFillStride};
  const PartSet fine0{fh, nullptr, ffill, cap_f, nf, 1, 1};
  const PartSet fine{fh, fa, ffill, cap_f, nf, 1, 1};
  const PartSide side{sh, ctr, ctr + 1, side_cap};
  const PartSide acc{ah, ctr + 2, ctr + 3, side_cap};
  hipStream_t s = ctx->stream;
  std::vector<hipEvent_t> ev(10, nullptr);
  for (auto& e : ev) HIPCHK(ctx, hipEventCreate(&e));
  struct EvFree {
    std::vector<hipEvent_t>& v;
    ~EvFree() {
      for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    }
  } ev_free{ev};
  auto zero = [&](u32* p, u64 words) { return hipMemsetAsync(p, 0, words * 4, s) == hipSuccess; };
  bool ok = launch_part_low(g, lo, s) == 0 && launch_part_gen(x, n, 7, flat, s) == 0 &&
            zero(ctr, 4) && hipMemsetAsync(sums, 0, 32, s) == hipSuccess;
  // the timed pipeline (events between the stages)
  ok = ok && hipEventRecord(ev[0], s) == hipSuccess && zero(cfill, ncr * kCoarseFillStride) &&
       launch_part_first(flat, n, nullptr, g, coarse0, side, s) == 0 && hipEventRecord(ev[1], s) == hipSuccess;
  for (int p = 0; p < 3 && ok; ++p) {
    const PartSet& cin = p == 0 ? coarse0 : coarse;
    const PartSet& fout = p == 0 ? fine0 : fine;
    ok = zero(ffill, nf) && launch_part_fine(p, g, cin, fout, side, s) == 0 &&
         hipEventRecord(ev[2 + 2 * p], s) == hipSuccess;
    if (ok && p < 2)
      ok = zero(cfill, ncr * kCoarseFillStride) && launch_part_lookup(p, g, fout, coarse, side, s) == 0;
    else if (ok)
      ok = launch_part_last(g, fine, side, s) == 0;
    ok = ok && hipEventRecord(ev[3 + 2 * p], s) == hipSuccess;
  }
  ok = ok && launch_part_check(g, side, acc, ctx->num_cu * 4, s) == 0 && hipEventRecord(ev[8], s) == hipSuccess;
  // direct queries of the same hashes
  ok = ok && launch_part_ref(x, flat, n, sums + 2, ctx->num_cu * 8, s) == 0 &&
       hipEventRecord(ev[9], s) == hipSuccess && launch_part_sum(ah, ctr + 2, side_cap, sums, s) == 0 &&
       hipStreamSynchronize(s) == hipSuccess;
  if (!ok) FAIL(ctx, FC5_E_HIP, "partitioned probing measurement failed");
  float t[9] = {0};
  for (int i = 0; i < 9; ++i) HIPCHK(ctx, hipEventElapsedTime(&t[i], ev[i], ev[i + 1]));
  // ms: [0] first partition, [1] fine 0, [2] lookup 0, [3] fine 1, [4] lookup 1,
  // [5] fine 2, [6] last lookup, [7] full queries, [8] direct queries
  for (int i = 0; i < 9; ++i) ms[i] = t[i];
  u32 c[4];
  unsigned long long sm[4];
  HIPCHK(ctx, hipMemcpy(c, ctr, sizeof c, hipMemcpyDeviceToHost));
  HIPCHK(ctx, hipMemcpy(sm, sums, sizeof sm, hipMemcpyDeviceToHost));
  // counts: [0] accepted, [1] their hash sum, [2] accepted by direct queries,
  // [3] their hash sum, [4] full queries (candidates + overflow), [5] lost
  // (side or accepted list full), [6] slices, [7] coarse buckets
  counts[0] = sm[0];
  counts[1] = sm[1];
  counts[2] = sm[2];
  counts[3] = sm[3];
  counts[4] = c[0];
  counts[5] = (u64)c[1] + c[3];
  counts[6] = ns;
  counts[7] = g.nc;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}
