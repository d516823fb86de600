// fc5_strict.hpp -- the strict residue gate (option "strict_gate").
//
// The reference gate keeps T when T mod 504*247 is a sum of two sixth powers
// (fourcore_cls5_gpu_v4.cu:124-137).  The strict gate adds the same necessary
// condition modulo 64, 81 and 31*37*43: every T = a^6 + b^6 (a, b >= 0) passes
// it, so no solution is lost, and only the xor filter's false positives among
// the rejected T go missing from the maybe-hit list.  On the reference gate's
// survivors of the B ~ 3.2M class-5 band about 9.6% pass.
//
// The kernels test it on T = X - L with X the row part (one value per e-row of
// a class-5 unit, or per d-row of a class 2-4 unit, known exactly in 128 bits)
// and L the lane part (floor(d^6/64) for class 5, c^6 for classes 2-4):
//   row word   X mod 31*37*43 (bits 0-15), X mod 81 (bits 16-22), X mod 64 (bits 23-28)
//   lane part  L mod 31*37*43, L mod 81, L mod 64, all in 32-bit arithmetic
#pragma once

#include <cstdint>

#include "fc5_moddiv.hpp"  // FC5_HD

namespace fc5 {

constexpr uint32_t kStrictA = 31u * 37u * 43u;  // 49321: products of residues fit 32 bits

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
constexpr uint32_t kInv64A = inv_mod_small(64, kStrictA);
constexpr uint32_t kInv64_81 = inv_mod_small(64, 81);
static_assert(kInv64A != 0 && kInv64_81 != 0, "64 is invertible modulo 31*37*43 and 81");
static_assert((uint64_t)kStrictA * kStrictA < (1ULL << 32), "residue products fit 32 bits");

FC5_HD uint32_t bit_of(uint64_t m, uint32_t s) { return (uint32_t)(m >> s) & 1u; }

// the predicate on T's residues: ta = T mod 31*37*43, t81 = T mod 81, t64 = T mod 64
FC5_HD bool strict_pass_t(uint32_t ta, uint32_t t81, uint32_t t64) {
  const uint32_t b81 = t81 < 64u ? bit_of(kSum81Lo, t81) : bit_of(kSum81Hi, t81 - 64u);
  return (bit_of(kSum64, t64) & b81 & bit_of(kSum31, ta % 31u) & bit_of(kSum37, ta % 37u) &
          bit_of(kSum43, ta % 43u)) != 0;
}

// X mod m for the 128-bit X = xh:xl (m < 2^16)
FC5_HD uint32_t mod128_small(uint64_t xh, uint64_t xl, uint32_t m) {
  const uint64_t two64 = (~0ULL % m + 1) % m;
  return (uint32_t)(((xh % m) * two64 + xl % m) % m);
}

// the row word of X = xh:xl
FC5_HD uint32_t strict_row(uint64_t xh, uint64_t xl) {
  return mod128_small(xh, xl, kStrictA) | mod128_small(xh, xl, 81u) << 16 | (uint32_t)(xl & 63u) << 23;
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

// lane part L = floor(x^6/64) (class 5, x = d): L = (x^6 - rho) / 64 with rho = x^6 mod 64
FC5_HD void strict_lane_q64(uint32_t x, uint32_t& la, uint32_t& l81, uint32_t& l64) {
  const uint32_t r6 = pow6_4096(x), rho = r6 & 63u;
  l64 = r6 >> 6;
  uint32_t s = pow6_mod16(x, kStrictA);
  s = s >= rho ? s - rho : s + kStrictA - rho;
  la = s * kInv64A % kStrictA;
  uint32_t t = pow6_mod16(x, 81u);
  t = t >= rho ? t - rho : t + 81u - rho;
  l81 = t * kInv64_81 % 81u;
}

// lane part L = x^6 (classes 2-4, x = c3)
FC5_HD void strict_lane_p6(uint32_t x, uint32_t& la, uint32_t& l81, uint32_t& l64) {
  la = pow6_mod16(x, kStrictA);
  l81 = pow6_mod16(x, 81u);
  l64 = pow6_4096(x) & 63u;
}

// T = X - L in the strict residue sets, from X's row word and L's lane part
FC5_HD bool strict_pass(uint32_t row, uint32_t la, uint32_t l81, uint32_t l64) {
  uint32_t ta = (row & 0xFFFFu) + kStrictA - la;
  if (ta >= kStrictA) ta -= kStrictA;
  uint32_t t81 = ((row >> 16) & 127u) + 81u - l81;
  if (t81 >= 81u) t81 -= 81u;
  return strict_pass_t(ta, t81, ((row >> 23) - l64) & 63u);
}

}  // namespace fc5
