// fc5_moddiv.hpp -- exact h % d for a fixed divisor d (host and device).
//
// For d < 2^32: the 2/1 division with a precomputed reciprocal of Moller and
// Granlund ("Improved division by invariant integers", IEEE Trans. Computers
// 60(2), 2011, Algorithm 4), run on 32-bit words over the normalised
// numerator h << s (three words).  For d >= 2^32 the plain operator.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define FC5_HD __host__ __device__ __forceinline__
#else
#define FC5_HD inline
#endif

namespace fc5 {

struct ModDiv {
  std::uint64_t d;
  std::uint32_t dn;   // d << s (normalised, top bit set)
  std::uint32_t v;    // floor((2^64-1)/dn) - 2^32
  std::uint32_t s;    // shift
  std::uint32_t big;  // d >= 2^32 (or d == 0)
};

inline ModDiv make_moddiv(std::uint64_t d) {
  ModDiv m{};
  m.d = d;
  if (d == 0 || d >= (1ULL << 32)) {
    m.big = 1;
    return m;
  }
  m.s = (std::uint32_t)__builtin_clz((std::uint32_t)d);
  m.dn = (std::uint32_t)d << m.s;
  m.v = (std::uint32_t)(~0ULL / m.dn - (1ULL << 32));
  return m;
}

// remainder of (u1*2^32 + u0) / dn for u1 < dn, dn normalised
FC5_HD std::uint32_t mg_step(std::uint32_t u1, std::uint32_t u0, std::uint32_t dn, std::uint32_t v) {
  const std::uint64_t q = (std::uint64_t)v * u1 + (((std::uint64_t)u1 << 32) | u0);
  const std::uint32_t q1 = (std::uint32_t)(q >> 32) + 1u;
  const std::uint32_t q0 = (std::uint32_t)q;
  std::uint32_t r = u0 - q1 * dn;
  if (r > q0) r += dn;
  if (r >= dn) r -= dn;
  return r;
}

FC5_HD std::uint64_t mod_fast(std::uint64_t h, const ModDiv& m) {
  if (m.big) return h % m.d;
  const std::uint32_t s = m.s;
  const std::uint32_t n2 = s ? (std::uint32_t)(h >> (64 - s)) : 0u;
  const std::uint64_t hs = h << s;
  std::uint32_t r = mg_step(n2, (std::uint32_t)(hs >> 32), m.dn, m.v);
  r = mg_step(r, (std::uint32_t)hs, m.dn, m.v);
  return (std::uint64_t)(r >> s);
}

}  // namespace fc5
