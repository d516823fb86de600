// fc5_stage1_api.cpp -- Stage-1 on the device: fc5_emit_units_device,
// fc5_band_create_range and the emission helper used by fc5_serve_submit_range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "fc5_ctx.hpp"

using namespace fc5;

namespace {

struct Stage1Tables {
  std::vector<u64> z2, z3, z7;  // sixth roots of unity (units) mod 2^6, 3^6, 7^6
  u64 e2_14 = 0, e7_14 = 0, e3_21 = 0, e7_21 = 0;
};

u64 inv_mod_u64(u64 a, u64 m) {  // a coprime to m
  long long t = 0, nt = 1, r = (long long)m, nr = (long long)(a % m);
  while (nr) {
    const long long q = r / nr;
    long long x = t - q * nt;
    t = nt;
    nt = x;
    x = r - q * nr;
    r = nr;
    nr = x;
  }
  return (u64)(t < 0 ? t + (long long)m : t);
}

const Stage1Tables& tables() {
  static const Stage1Tables t = [] {
    Stage1Tables s;
    auto roots = [](u64 m, u64 p) {
      std::vector<u64> z;
      for (u64 x = 1; x < m; ++x)
        if (x % p && pow6_mod((i64)x, (i64)m) == 1) z.push_back(x);
      return z;
    };
    s.z2 = roots(64, 2);
    s.z3 = roots(729, 3);
    s.z7 = roots(117649, 7);
    s.e2_14 = 117649ull * inv_mod_u64(117649 % 64, 64);         // 1 mod 2^6, 0 mod 7^6
    s.e7_14 = 64ull * inv_mod_u64(64, 117649);                  // 0 mod 2^6, 1 mod 7^6
    s.e3_21 = 117649ull * inv_mod_u64(117649 % 729, 729);       // 1 mod 3^6, 0 mod 7^6
    s.e7_21 = 729ull * inv_mod_u64(729, 117649);                // 0 mod 3^6, 1 mod 7^6
    return s;
  }();
  return t;
}

constexpr u64 kChunkB = 1ull << 21;  // B values per launch pair

}  // namespace

int fc5::stage1_emit(fc5_ctx* ctx, hipStream_t s, u64 lo, u64 hi, unsigned mask, double lo_frac, double hi_frac,
                     std::vector<u32>& cls, std::vector<fc5_unit>& units) {
  cls.clear();
  units.clear();
  if (lo == 0 || hi < lo) FAIL(ctx, FC5_E_ARG, "bad B range");
  if (mask & ~0x3cu) FAIL(ctx, FC5_E_ARG, "device Stage-1 emits classes 2-5 (class 1 is host-only)");
  const Stage1Tables& T = tables();
  if (T.z2.size() != 4 || T.z3.size() != 6 || T.z7.size() != 6) FAIL(ctx, FC5_E_STATE, "roots of unity");
  u64* d_z = nullptr;
  u64 *d_counts = nullptr, *d_offsets = nullptr;
  void* d_temp = nullptr;
  u32* d_cls = nullptr;
  u64 *d_B = nullptr, *d_u = nullptr;
  size_t out_cap = 0, temp_bytes = 0;
  auto release = [&] {
    for (void* p : {(void*)d_z, (void*)d_counts, (void*)d_offsets, d_temp, (void*)d_cls, (void*)d_B, (void*)d_u})
      if (p) (void)hipFree(p);
  };
  int rc = FC5_OK;
  const u64 chunk = std::min<u64>(kChunkB, hi - lo + 1);
  std::vector<u64> z(16);
  std::copy(T.z2.begin(), T.z2.end(), z.begin());
  std::copy(T.z3.begin(), T.z3.end(), z.begin() + 4);
  std::copy(T.z7.begin(), T.z7.end(), z.begin() + 10);
  if (hipMalloc(&d_z, 16 * 8) != hipSuccess || hipMalloc(&d_counts, chunk * 8) != hipSuccess ||
      hipMalloc(&d_offsets, chunk * 8) != hipSuccess ||
      stage1_scan(d_counts, d_offsets, chunk, nullptr, temp_bytes, s) != 0 ||
      hipMalloc(&d_temp, std::max<size_t>(temp_bytes, 8)) != hipSuccess ||
      hipMemcpy(d_z, z.data(), 16 * 8, hipMemcpyHostToDevice) != hipSuccess) {
    release();
    FAIL(ctx, FC5_E_NOMEM, "Stage-1 buffers");
  }
  Stage1Args a{};
  a.lo_frac = lo_frac;
  a.hi_frac = hi_frac;
  a.cls_mask = mask;
  a.z2 = d_z;
  a.z3 = d_z + 4;
  a.z7 = d_z + 10;
  a.e2_14 = T.e2_14;
  a.e7_14 = T.e7_14;
  a.e3_21 = T.e3_21;
  a.e7_21 = T.e7_21;
  a.counts = d_counts;
  a.offsets = d_offsets;
  for (u64 b0 = lo; b0 <= hi && rc == FC5_OK; b0 += chunk) {
    a.lo = b0;
    a.count = std::min<u64>(chunk, hi - b0 + 1);
    u64 last_off = 0, last_cnt = 0;
    size_t tb = temp_bytes;
    if (launch_stage1(a, 0, s) != 0 || stage1_scan(d_counts, d_offsets, a.count, d_temp, tb, s) != 0 ||
        hipMemcpyAsync(&last_off, d_offsets + a.count - 1, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&last_cnt, d_counts + a.count - 1, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      ctx->err = "Stage-1 count pass";
      rc = FC5_E_HIP;
      break;
    }
    const u64 total = last_off + last_cnt;
    if (!total) continue;
    if (total > out_cap) {
      for (void* p : {(void*)d_cls, (void*)d_B, (void*)d_u})
        if (p) (void)hipFree(p);
      d_cls = nullptr;
      d_B = d_u = nullptr;
      out_cap = total;
      if (hipMalloc(&d_cls, out_cap * 4) != hipSuccess || hipMalloc(&d_B, out_cap * 8) != hipSuccess ||
          hipMalloc(&d_u, out_cap * 8) != hipSuccess) {
        ctx->err = "Stage-1 output buffers";
        rc = FC5_E_NOMEM;
        break;
      }
    }
    a.out_cls = d_cls;
    a.out_B = d_B;
    a.out_u = d_u;
    const size_t at = units.size();
    std::vector<u32> hc(total);
    std::vector<u64> hb(total), hu(total);
    if (launch_stage1(a, 1, s) != 0 ||
        hipMemcpyAsync(hc.data(), d_cls, total * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hb.data(), d_B, total * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hu.data(), d_u, total * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      ctx->err = "Stage-1 write pass";
      rc = FC5_E_HIP;
      break;
    }
    cls.insert(cls.end(), hc.begin(), hc.end());
    units.resize(at + total);
    for (u64 i = 0; i < total; ++i) units[at + i] = fc5_unit{hb[i], hu[i]};
  }
  release();
  return rc;
}

extern "C" int fc5_emit_units_device(fc5_ctx* ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac,
                                     double hi_frac, uint32_t* out_cls, uint64_t* out_B, uint64_t* out_u, size_t cap,
                                     size_t* n) try {
  if (!ctx || !n || (cap && (!out_cls || !out_B || !out_u))) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<u32> cls;
  std::vector<fc5_unit> units;
  const int rc = stage1_emit(ctx, ctx->stream, lo, hi, cls_mask, lo_frac, hi_frac, cls, units);
  if (rc != FC5_OK) return rc;
  *n = units.size();
  for (size_t i = 0; i < units.size() && i < cap; ++i) {
    out_cls[i] = cls[i];
    out_B[i] = units[i].B;
    out_u[i] = units[i].u;
  }
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_band_create_range(fc5_ctx* ctx, uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac,
                                     double hi_frac, fc5_band** out) try {
  if (!ctx || !out) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<u32> cls;
  std::vector<fc5_unit> units;
  const int rc = stage1_emit(ctx, ctx->stream, lo, hi, cls_mask, lo_frac, hi_frac, cls, units);
  if (rc != FC5_OK) return rc;
  return fc5_band_create_mixed(ctx, cls.data(), units.data(), units.size(), out);
} catch (...) {
  return abi_exception(ctx);
}
