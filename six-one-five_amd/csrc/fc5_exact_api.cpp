// fc5_exact_api.cpp -- the exact sorted pair table (QuadSumIndex of
// quad_sum.hpp:8-45, built and kept on the device): build, read-back, lookups,
// and the exact-table half of the hit verification (device_pairs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "fc5_ctx.hpp"

using namespace fc5;

namespace {

u32 bit_length(u128 v) {
  u32 n = 0;
  while (v) {
    ++n;
    v >>= 1;
  }
  return n;
}

constexpr int kExactMaxN = 262143;  // 2 N^6 * 2^18 < 2^128

}  // namespace

void fc5::free_exact(fc5_ctx* ctx) {
  if (ctx->d_exact) (void)hipFree(ctx->d_exact);
  ctx->d_exact = nullptr;
  ctx->n_exact = 0;
  ctx->exact_N = 0;
  ctx->exact_shift = 0;
}

extern "C" int fc5_exact_build(fc5_ctx* ctx, int N, const uint64_t* extra, size_t n_extra) try {
  if (!ctx) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (N < 1 || N > kExactMaxN) FAIL(ctx, FC5_E_ARG, "exact table N must be 1..262143");
  if (n_extra && !extra) FAIL(ctx, FC5_E_ARG, "extra values missing");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const u32 s = bit_length((u128)N);
  const u64 n3 = (u64)N * N * N;
  u128 maxv = 2 * ((u128)n3 * n3);
  std::vector<u128> ex(n_extra);
  for (size_t i = 0; i < n_extra; ++i) {
    ex[i] = ((u128)extra[2 * i + 1] << 64) | extra[2 * i];
    if (ex[i] == 0) FAIL(ctx, FC5_E_ARG, "extra values must be positive");
    maxv = std::max(maxv, ex[i]);
  }
  const u32 end_bit = bit_length(maxv) + s;
  if (end_bit > 128) FAIL(ctx, FC5_E_ARG, "an extra value is too large: value * 2^s + a must fit 128 bits");
  for (u128& v : ex) v <<= s;
  const u64 n_pairs = (u64)N * ((u64)N + 1) / 2, n = n_pairs + n_extra;
  free_exact(ctx);
  u128 *keys = nullptr, *alt = nullptr;
  void* temp = nullptr;
  auto release = [&] {
    for (void* p : {(void*)keys, (void*)alt, temp})
      if (p) (void)hipFree(p);
  };
  size_t temp_bytes = 0;
  bool in_alt = false;
  if (exact_sort(nullptr, nullptr, n, end_bit, nullptr, temp_bytes, ctx->stream, in_alt) != 0)
    FAIL(ctx, FC5_E_HIP, "exact table: sort size query");
  if (hipMalloc(&keys, n * sizeof(u128)) != hipSuccess || hipMalloc(&alt, n * sizeof(u128)) != hipSuccess ||
      hipMalloc(&temp, std::max<size_t>(temp_bytes, 16)) != hipSuccess) {
    release();
    FAIL(ctx, FC5_E_NOMEM, "exact table: device memory for " + std::to_string(n) + " entries (32 B each while sorting)");
  }
  if (launch_exact_rows(keys, (u32)N, s, ctx->stream) != 0 ||
      (n_extra && hipMemcpyAsync(keys + n_pairs, ex.data(), n_extra * sizeof(u128), hipMemcpyHostToDevice,
                                 ctx->stream) != hipSuccess) ||
      exact_sort(keys, alt, n, end_bit, temp, temp_bytes, ctx->stream, in_alt) != 0 ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    release();
    FAIL(ctx, FC5_E_HIP, "exact table build");
  }
  if (in_alt) std::swap(keys, alt);
  (void)hipFree(alt);
  (void)hipFree(temp);
  ctx->d_exact = keys;
  ctx->n_exact = n;
  ctx->exact_N = N;
  ctx->exact_shift = s;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_exact_clear(fc5_ctx* ctx) try {
  if (!ctx) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_exact(ctx);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_exact_info(const fc5_ctx* ctx, uint64_t* n_entries, int* N, uint32_t* shift) {
  if (!ctx) return FC5_E_ARG;
  if (n_entries) *n_entries = ctx->n_exact;
  if (N) *N = ctx->exact_N;
  if (shift) *shift = ctx->exact_shift;
  return FC5_OK;
}

extern "C" int fc5_exact_read(fc5_ctx* ctx, uint64_t first, size_t n, uint64_t* out) try {
  if (!ctx || (n && !out)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (first > ctx->n_exact || n > ctx->n_exact - first) FAIL(ctx, FC5_E_ARG, "entries out of range");
  if (!n) return FC5_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipMemcpy(out, ctx->d_exact + first, n * sizeof(u128), hipMemcpyDeviceToHost));
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

int fc5::exact_table_pairs(fc5_ctx* ctx, const std::vector<PairTarget>& targets,
                           std::vector<std::vector<std::pair<u32, u32>>>& pairs, std::vector<u8>* extra,
                           hipStream_t st) {
  pairs.assign(targets.size(), {});
  if (extra) extra->assign(targets.size(), 0);
  if (!ctx->d_exact) FAIL(ctx, FC5_E_STATE, "verify_table is set but no exact table is built");
  std::vector<PairJob> jobs;
  std::vector<size_t> which;
  for (size_t i = 0; i < targets.size(); ++i) {
    if (!targets[i].ok) continue;
    if (targets[i].hi > (u32)ctx->exact_N)
      FAIL(ctx, FC5_E_STATE, "exact table N=" + std::to_string(ctx->exact_N) + " is below the pair bound " +
                                 std::to_string(targets[i].hi));
    jobs.push_back(PairJob{(u64)targets[i].T, (u64)(targets[i].T >> 64), targets[i].hi, 0});
    which.push_back(i);
  }
  const u128* keys = ctx->d_exact;
  const u64 n = ctx->n_exact;
  const u32 s = ctx->exact_shift;
  std::vector<PairHit> out;
  const int rc = run_pair_jobs(
      ctx, jobs,
      [keys, n, s](const PairJob* j, u32 nj, PairHit* o, u32 cap, u32* count, hipStream_t st) {
        return launch_exact_lookup(keys, n, s, j, nj, o, cap, count, st);
      },
      out, st);
  if (rc != FC5_OK) return rc;
  for (const PairHit& h : out) {
    if (h.pad) {
      if (extra) (*extra)[which[h.job]] = 1;
    } else {
      pairs[which[h.job]].emplace_back(h.a, h.b);
    }
  }
  for (auto& v : pairs) std::sort(v.begin(), v.end());
  return FC5_OK;
}

extern "C" int fc5_exact_lookup(fc5_ctx* ctx, const uint64_t* targets, const uint32_t* hi, size_t n, uint32_t* out,
                                size_t cap, size_t* n_out) try {
  if (!ctx || !n_out || (n && (!targets || !hi)) || (cap && !out)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<PairTarget> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = PairTarget{((u128)targets[2 * i + 1] << 64) | targets[2 * i], hi[i], true};
  std::vector<std::vector<std::pair<u32, u32>>> pairs;
  std::vector<u8> extra;
  const int rc = exact_table_pairs(ctx, t, pairs, &extra);
  if (rc != FC5_OK) return rc;
  size_t k = 0;
  auto put = [&](size_t i, u32 a, u32 b) {
    if (k < cap) {
      out[3 * k] = (uint32_t)i;
      out[3 * k + 1] = a;
      out[3 * k + 2] = b;
    }
    ++k;
  };
  for (size_t i = 0; i < n; ++i) {
    if (extra[i]) put(i, 0, 0);
    for (auto [a, b] : pairs[i]) put(i, a, b);
  }
  *n_out = k;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}
