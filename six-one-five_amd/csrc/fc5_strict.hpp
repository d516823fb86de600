// fc5_strict.hpp -- the strict residue gate (option "strict_gate").
//
// The reference gate keeps T when T mod 504*247 is a sum of two sixth powers
// (fourcore_cls5_gpu_v4.cu:124-137).  The strict gate adds the same necessary
// condition modulo 64, 81, 31, 37, 43, 49, 61, 67, 73 and 79: every
// T = a^6 + b^6 (a, b >= 0) passes it, so no solution is lost, and only the
// xor filter's false positives among the rejected T go missing from the
// maybe-hit list.  (Every residue mod a prime p = 5 mod 6 is a sum of two
// sixth powers, so only p = 1 mod 6 and powers of 2, 3 and 7 can filter; on
// the reference gate's survivors of the B ~ 3.2M class-5 band these ten keep
// about 3%.)
//
// The kernels test it on T = X - L with X the row part (one value per e-row of
// a class-5 unit, or per d-row of a class 2-4 unit, known exactly in 128 bits)
// and L the lane part (floor(d^6/64) for class 5, c^6 for classes 2-4):
//   row words  w0: X mod 31*37*43 (bits 0-15), X mod 81 (16-22), X mod 64 (23-28)
//              w1: X mod 49*61*67 (bits 0-17), X mod 73*79 (18-30)
//   lane part  L mod each of those products, in 32/64-bit arithmetic
#pragma once

#include <cstdint>

#include "fc5_moddiv.hpp"  // FC5_HD

namespace fc5 {

constexpr uint32_t kStrictA = 31u * 37u * 43u;  // 49321: products of residues fit 32 bits
constexpr uint32_t kStrictB = 49u * 61u * 67u;  // 200263: products of residues in 64 bits
constexpr uint32_t kStrictC = 73u * 79u;        // 5767

// bit s of the mask: s mod m is a sum of two sixth powers mod m (m <= 128)
struct SumMask {
  uint64_t lo, hi;
};

constexpr SumMask sum2_pow6_mask(uint32_t m) {
  bool is6[128] = {};
  for (uint32_t x = 0; x < m; ++x) {
    uint64_t p = 1;
    for (int k = 0; k < 6; ++k) p = p * x % m;
    is6[p] = true;
  }
  SumMask r{0, 0};
  for (uint32_t a = 0; a < m; ++a)
    for (uint32_t b = 0; b < m; ++b)
      if (is6[a] && is6[b]) {
        const uint32_t s = (a + b) % m;
        if (s < 64) r.lo |= 1ULL << s;
        else r.hi |= 1ULL << (s - 64);
      }
  return r;
}

constexpr uint32_t inv_mod_small(uint32_t a, uint32_t m) {
  for (uint32_t x = 1; x < m; ++x)
    if ((uint64_t)a * x % m == 1) return x;
  return 0;
}

constexpr uint64_t kSum31 = sum2_pow6_mask(31).lo;
constexpr uint64_t kSum37 = sum2_pow6_mask(37).lo;
constexpr uint64_t kSum43 = sum2_pow6_mask(43).lo;
constexpr uint64_t kSum64 = sum2_pow6_mask(64).lo;
constexpr uint64_t kSum81Lo = sum2_pow6_mask(81).lo;
constexpr uint64_t kSum81Hi = sum2_pow6_mask(81).hi;
constexpr uint64_t kSum49 = sum2_pow6_mask(49).lo;
constexpr uint64_t kSum61 = sum2_pow6_mask(61).lo;
constexpr uint64_t kSum67Lo = sum2_pow6_mask(67).lo, kSum67Hi = sum2_pow6_mask(67).hi;
constexpr uint64_t kSum73Lo = sum2_pow6_mask(73).lo, kSum73Hi = sum2_pow6_mask(73).hi;
constexpr uint64_t kSum79Lo = sum2_pow6_mask(79).lo, kSum79Hi = sum2_pow6_mask(79).hi;
constexpr uint32_t kInv64A = inv_mod_small(64, kStrictA);
constexpr uint32_t kInv64_81 = inv_mod_small(64, 81);
constexpr uint32_t kInv64B = inv_mod_small(64, kStrictB);
constexpr uint32_t kInv64C = inv_mod_small(64, kStrictC);
static_assert(kInv64A != 0 && kInv64_81 != 0 && kInv64B != 0 && kInv64C != 0,
              "64 is invertible modulo 31*37*43, 81, 49*61*67 and 73*79");
static_assert((uint64_t)kStrictA * kStrictA < (1ULL << 32), "residue products fit 32 bits");
static_assert((uint64_t)kStrictC * kStrictC < (1ULL << 32), "residue products fit 32 bits");

FC5_HD uint32_t bit_of(uint64_t m, uint32_t s) { return (uint32_t)(m >> s) & 1u; }
FC5_HD uint32_t bit_of(uint64_t lo, uint64_t hi, uint32_t s) { return s < 64u ? bit_of(lo, s) : bit_of(hi, s - 64u); }

// the predicate on T's residues: ta = T mod 31*37*43, t81 = T mod 81, t64 = T mod 64,
// tb = T mod 49*61*67, tc = T mod 73*79
FC5_HD bool strict_pass_t(uint32_t ta, uint32_t t81, uint32_t t64, uint32_t tb, uint32_t tc) {
  return (bit_of(kSum64, t64) & bit_of(kSum81Lo, kSum81Hi, t81) & bit_of(kSum31, ta % 31u) &
          bit_of(kSum37, ta % 37u) & bit_of(kSum43, ta % 43u) & bit_of(kSum49, tb % 49u) &
          bit_of(kSum61, tb % 61u) & bit_of(kSum67Lo, kSum67Hi, tb % 67u) & bit_of(kSum73Lo, kSum73Hi, tc % 73u) &
          bit_of(kSum79Lo, kSum79Hi, tc % 79u)) != 0;
}

// the row part's two words and a lane part's residues
struct StrictRow {
  uint32_t w0, w1;
};
struct StrictLane {
  uint32_t la, l81, l64, lb, lc;
};

// X mod m for the 128-bit X = xh:xl (m < 2^31)
FC5_HD uint32_t mod128_small(uint64_t xh, uint64_t xl, uint32_t m) {
  const uint64_t two64 = (~0ULL % m + 1) % m;
  return (uint32_t)(((xh % m) * two64 + xl % m) % m);
}

// the row words of X = xh:xl
FC5_HD StrictRow strict_row(uint64_t xh, uint64_t xl) {
  return {mod128_small(xh, xl, kStrictA) | mod128_small(xh, xl, 81u) << 16 | (uint32_t)(xl & 63u) << 23,
          mod128_small(xh, xl, kStrictB) | mod128_small(xh, xl, kStrictC) << 18};
}

// x^6 mod m for m < 2^18 (64-bit products)
FC5_HD uint32_t pow6_mod18(uint32_t x, uint32_t m) {
  const uint64_t r = x % m, r2 = r * r % m, r3 = r2 * r % m;
  return (uint32_t)(r3 * r3 % m);
}

// x^6 mod m for m < 2^16
FC5_HD uint32_t pow6_mod16(uint32_t x, uint32_t m) {
  const uint32_t r = x % m, r2 = r * r % m, r3 = r2 * r % m;
  return r3 * r3 % m;
}

// x^6 mod 4096
FC5_HD uint32_t pow6_4096(uint32_t x) {
  const uint32_t r = x & 4095u, r2 = (r * r) & 4095u, r3 = (r2 * r) & 4095u;
  return (r3 * r3) & 4095u;
}

// (s - rho) / 64 mod m for s = x^6 mod m, rho = x^6 mod 64 (m odd)
FC5_HD uint32_t div64_mod(uint32_t s, uint32_t rho, uint32_t m, uint32_t inv64) {
  const uint32_t t = s >= rho % m ? s - rho % m : s + m - rho % m;
  return (uint32_t)((uint64_t)t * inv64 % m);
}

// lane part L = floor(x^6/64) (class 5, x = d): L = (x^6 - rho) / 64 with rho = x^6 mod 64
FC5_HD StrictLane strict_lane_q64(uint32_t x) {
  const uint32_t r6 = pow6_4096(x), rho = r6 & 63u;
  StrictLane L;
  L.l64 = r6 >> 6;
  L.la = div64_mod(pow6_mod16(x, kStrictA), rho, kStrictA, kInv64A);
  L.l81 = div64_mod(pow6_mod16(x, 81u), rho, 81u, kInv64_81);
  L.lb = div64_mod(pow6_mod18(x, kStrictB), rho, kStrictB, kInv64B);
  L.lc = div64_mod(pow6_mod16(x, kStrictC), rho, kStrictC, kInv64C);
  return L;
}

// lane part L = x^6 (classes 2-4, x = c3)
FC5_HD StrictLane strict_lane_p6(uint32_t x) {
  StrictLane L;
  L.la = pow6_mod16(x, kStrictA);
  L.l81 = pow6_mod16(x, 81u);
  L.l64 = pow6_4096(x) & 63u;
  L.lb = pow6_mod18(x, kStrictB);
  L.lc = pow6_mod16(x, kStrictC);
  return L;
}

FC5_HD uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t m) { return a >= b ? a - b : a + m - b; }

// T = X - L in the strict residue sets, from X's row words and L's lane part
FC5_HD bool strict_pass(const StrictRow& row, const StrictLane& L) {
  return strict_pass_t(sub_mod(row.w0 & 0xFFFFu, L.la, kStrictA), sub_mod((row.w0 >> 16) & 127u, L.l81, 81u),
                       ((row.w0 >> 23) - L.l64) & 63u, sub_mod(row.w1 & 0x3FFFFu, L.lb, kStrictB),
                       sub_mod(row.w1 >> 18, L.lc, kStrictC));
}

}  // namespace fc5
