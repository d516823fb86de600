// fc5_sieve.hpp -- the residue gate of the class-5 sweep as a bit sieve over d.
//
// The reference gate keeps T when T mod 504 and T mod 247 are sums of two
// sixth powers (fourcore_cls5_gpu_v4.cu:124-137,393-402).  A residue is a sum
// of two sixth powers mod 504 = 8*9*7 iff it is one mod 8, mod 9 and mod 7
// (by the Chinese remainder theorem the two bases range over all residues
// modulo each factor independently), and likewise mod 247 = 13*19, so the
// gate is the conjunction of five conditions on T mod 8, 9, 7, 13 and 19.
//
// In a class-5 row (one e) T = X - floor(d^6/64) over d = r + 64k, where r is
// one of the unit's four residues mod 64 and d^6 = td (mod 64) for all of them:
//   odd m   T = (X + td/64) - d^6/64 (mod m): the row phase A = X + td/64 mod m
//           and d mod m (d mod 3 for m = 9, where x^6 mod 9 depends on x mod 3);
//   m = 8   T mod 8 = X mod 8 - ((d^6 mod 512) >> 6), a function of d mod 256.
// A sieve word covers k = 32w .. 32w+31 of one residue, d = d0 + 64j with
// d0 = r + 2048w; its bit j for modulus m is fixed by (A, j0), j0 = d0 mod m
// (mod 3 for m = 9) or j0 = r for m = 8 (2048w = 0 mod 256).  The table holds
// one word per (m, A, j0): 1118 words.  AND-ing the five words of a row and a
// word gives the gate survivors among 32 grid points.
#pragma once

#include <cstdint>

#include "fc5_moddiv.hpp"  // FC5_HD
#include "fc5_strict.hpp"  // StrictRow

namespace fc5 {

constexpr uint32_t kSieveOff8 = 0;     // [A = X mod 8][r]: 8 x 64
constexpr uint32_t kSieveOff9 = 512;   // [A][d0 mod 3]: 9 x 3
constexpr uint32_t kSieveOff7 = 539;   // [A][d0 mod 7]: 7 x 7
constexpr uint32_t kSieveOff13 = 588;  // [A][d0 mod 13]: 13 x 13
constexpr uint32_t kSieveOff19 = 757;  // [A][d0 mod 19]: 19 x 19
constexpr uint32_t kSieveWords = 1118;

struct SieveTab {
  uint32_t w[kSieveWords];
};

// bit t: t mod m is a sum of two sixth powers mod m (m <= 32)
constexpr uint32_t sieve_sum_mask(uint32_t m) {
  uint32_t p6 = 0;
  for (uint32_t x = 0; x < m; ++x) {
    uint64_t p = 1;
    for (int i = 0; i < 6; ++i) p = p * x % m;
    p6 |= 1u << p;
  }
  uint32_t s = 0;
  for (uint32_t a = 0; a < m; ++a)
    for (uint32_t b = 0; b < m; ++b)
      if (((p6 >> a) & 1u) && ((p6 >> b) & 1u)) s |= 1u << ((a + b) % m);
  return s;
}

constexpr uint32_t sieve_inv64(uint32_t m) {
  for (uint32_t x = 1; x < m; ++x)
    if (64u * x % m == 1u) return x;
  return 0;
}

constexpr uint32_t sieve_pow6(uint32_t x, uint32_t m) {
  uint64_t p = 1;
  for (int i = 0; i < 6; ++i) p = p * x % m;
  return (uint32_t)p;
}

constexpr SieveTab make_sieve() {
  SieveTab s{};
  const uint32_t g8 = sieve_sum_mask(8);
  for (uint32_t A = 0; A < 8; ++A)
    for (uint32_t r = 0; r < 64; ++r) {
      uint32_t word = 0;
      for (uint32_t j = 0; j < 32; ++j) {
        const uint32_t x = (r + 64u * j) & 255u;
        uint64_t x6 = 1;
        for (int i = 0; i < 6; ++i) x6 = x6 * x % 512u;
        const uint32_t dq8 = (uint32_t)(x6 >> 6);  // floor(d^6/64) mod 8
        if ((g8 >> ((A + 8u - dq8) & 7u)) & 1u) word |= 1u << j;
      }
      s.w[kSieveOff8 + A * 64u + r] = word;
    }
  const uint32_t mods[4] = {9, 7, 13, 19}, nj[4] = {3, 7, 13, 19};
  const uint32_t offs[4] = {kSieveOff9, kSieveOff7, kSieveOff13, kSieveOff19};
  for (int t = 0; t < 4; ++t) {
    const uint32_t m = mods[t], n = nj[t], g = sieve_sum_mask(m), inv = sieve_inv64(m);
    for (uint32_t A = 0; A < m; ++A)
      for (uint32_t j0 = 0; j0 < n; ++j0) {
        uint32_t word = 0;
        for (uint32_t j = 0; j < 32; ++j) {
          const uint32_t x = (j0 + 64u * j) % n;  // d mod n
          const uint32_t D = sieve_pow6(x, m) * inv % m;
          if ((g >> ((A + m - D) % m)) & 1u) word |= 1u << j;
        }
        s.w[offs[t] + A * n + j0] = word;
      }
  }
  return s;
}

static_assert(sieve_inv64(9) == 1 && sieve_inv64(7) == 1 && sieve_inv64(13) == 12 && sieve_inv64(19) == 11,
              "inverses of 64 used by sieve_row_bases");
static_assert(sieve_sum_mask(8) == 0x7u && sieve_sum_mask(9) == 0x7u && sieve_sum_mask(7) == 0x7u,
              "T mod 8, 9, 7 in {0, 1, 2}");

// The row's table bases from xg = X mod 124488 and td = d^6 mod 64:
// ra = b8 | b9 << 16, rb = b7 | b13 << 16, rc = b19 (each below kSieveWords).
FC5_HD void sieve_row_bases(uint32_t xg, uint32_t td, uint32_t& ra, uint32_t& rb, uint32_t& rc) {
  const uint32_t a8 = xg & 7u;
  const uint32_t a9 = (xg % 9u + td) % 9u;
  const uint32_t a7 = (xg % 7u + td) % 7u;
  const uint32_t a13 = (xg % 13u + 12u * td) % 13u;
  const uint32_t a19 = (xg % 19u + 11u * td) % 19u;
  ra = (kSieveOff8 + a8 * 64u) | (kSieveOff9 + a9 * 3u) << 16;
  rb = (kSieveOff7 + a7 * 7u) | (kSieveOff13 + a13 * 13u) << 16;
  rc = kSieveOff19 + a19 * 19u;
}

// The gate survivors among k = 32w .. 32w+31 of residue r (d0 = r + 2048w)
// before the cut, from the row's bases.
FC5_HD uint32_t sieve_word(const uint32_t* tab, uint32_t ra, uint32_t rb, uint32_t rc, uint32_t r, uint32_t d0) {
  return tab[(ra & 0xFFFFu) + r] & tab[(ra >> 16) + d0 % 3u] & tab[(rb & 0xFFFFu) + d0 % 7u] &
         tab[(rb >> 16) + d0 % 13u] & tab[rc + d0 % 19u];
}

// Grid points d0 + 64j <= dl of a word: bits 0 .. (dl - d0)/64 (none if dl < d0).
FC5_HD uint32_t sieve_cut(uint32_t d0, uint32_t dl) {
  if (dl < d0) return 0u;
  const uint32_t jm = (dl - d0) >> 6;
  return 0xFFFFFFFFu >> (31u - (jm < 31u ? jm : 31u));
}

// ---------------------------------------------------------------------------
// The strict residue gate (fc5_strict.hpp: T mod 64, 81, 31, 37, 43, 49, 61,
// 67, 73 and 79 sums of two sixth powers) as sieve words in the same form, for
// the class-5 rows:
//   m = 64  T mod 64 = X mod 64 - ((d^6 mod 4096) >> 6), a function of d mod
//           2048, so j0 = r (2048w = 0 mod 2048): [A = X mod 64][r]
//   m = 81  x^6 mod 81 depends on x mod 27: [A = X + td/64 mod 81][d0 mod 27]
//   other m [A = X + td/64 mod m][d0 mod m]
// 32643 words, read from global memory (the class-5 rows of the strict variant).
constexpr int kStrictMods = 10;
constexpr uint32_t kStrictMod[kStrictMods] = {64, 81, 31, 37, 43, 49, 61, 67, 73, 79};
constexpr uint32_t kStrictNj[kStrictMods] = {64, 27, 31, 37, 43, 49, 61, 67, 73, 79};  // j0 classes

constexpr uint32_t strict_sieve_off(int i) {
  uint32_t o = 0;
  for (int k = 0; k < i; ++k) o += kStrictMod[k] * kStrictNj[k];
  return o;
}
constexpr uint32_t kStrictWords = strict_sieve_off(kStrictMods);
static_assert(kStrictWords == 32643 && kStrictWords < (1u << 16), "table bases fit 16 bits");

// out[t]: t is a sum of two sixth powers mod m (m <= 128)
inline void sieve_sum_set(uint32_t m, bool* out) {
  bool p6[128] = {};
  for (uint32_t x = 0; x < m; ++x) p6[sieve_pow6(x, m)] = true;
  for (uint32_t t = 0; t < m; ++t) out[t] = false;
  for (uint32_t a = 0; a < m; ++a)
    for (uint32_t b = 0; b < m; ++b)
      if (p6[a] && p6[b]) out[(a + b) % m] = true;
}

inline void fill_strict_sieve(uint32_t* w) {
  bool S[128];
  sieve_sum_set(64, S);
  for (uint32_t A = 0; A < 64; ++A)
    for (uint32_t r = 0; r < 64; ++r) {
      uint32_t word = 0;
      for (uint32_t j = 0; j < 32; ++j) {
        const uint64_t x = r + 64u * j;  // d mod 2048
        uint64_t x6 = 1;
        for (int i = 0; i < 6; ++i) x6 = x6 * x % 4096u;
        if (S[(A + 64u - (uint32_t)(x6 >> 6)) & 63u]) word |= 1u << j;
      }
      w[A * 64u + r] = word;
    }
  for (int t = 1; t < kStrictMods; ++t) {
    const uint32_t m = kStrictMod[t], n = kStrictNj[t], inv = sieve_inv64(m), off = strict_sieve_off(t);
    sieve_sum_set(m, S);
    for (uint32_t A = 0; A < m; ++A)
      for (uint32_t j0 = 0; j0 < n; ++j0) {
        uint32_t word = 0;
        for (uint32_t j = 0; j < 32; ++j) {
          const uint32_t x = (j0 + 64u * j) % n;  // d mod n
          const uint32_t D = sieve_pow6(x, m) * inv % m;
          if (S[(A + m - D) % m]) word |= 1u << j;
        }
        w[off + A * n + j0] = word;
      }
  }
}

// a row's ten table bases, two to a word (b[2k] | b[2k+1] << 16), and a
// lane's nine j0 = d0 mod kStrictNj[1..9]
constexpr int kStrictRowWords = kStrictMods / 2;
struct StrictSieveRow {
  uint32_t w[kStrictRowWords];
};
struct StrictSieveLane {
  uint32_t j[kStrictMods - 1];
};

// The row's strict table bases from its strict row words (fc5_strict.hpp) and td.
FC5_HD StrictSieveRow strict_sieve_bases(const StrictRow& xs, uint32_t td) {
  const uint32_t xa = xs.w0 & 0xFFFFu, xb = xs.w1 & 0x3FFFFu, xc = xs.w1 >> 18;
  const uint32_t xm[kStrictMods] = {(xs.w0 >> 23) & 63u, (xs.w0 >> 16) & 127u, xa % 31u, xa % 37u, xa % 43u,
                                    xb % 49u, xb % 61u, xb % 67u, xc % 73u, xc % 79u};
  uint32_t b[kStrictMods];
  b[0] = xm[0] * 64u;
#pragma unroll
  for (int t = 1; t < kStrictMods; ++t) {
    const uint32_t m = kStrictMod[t];
    b[t] = strict_sieve_off(t) + (xm[t] + td * sieve_inv64(m)) % m * kStrictNj[t];
  }
  StrictSieveRow r;
#pragma unroll
  for (int k = 0; k < kStrictRowWords; ++k) r.w[k] = b[2 * k] | b[2 * k + 1] << 16;
  return r;
}

FC5_HD StrictSieveLane strict_sieve_lane(uint32_t d0) {
  StrictSieveLane l;
#pragma unroll
  for (int t = 1; t < kStrictMods; ++t) l.j[t - 1] = d0 % kStrictNj[t];
  return l;
}

// The strict survivors among a word's 32 grid points (before the cut).
FC5_HD uint32_t strict_sieve_word(const uint32_t* tab, const StrictSieveRow& b, const StrictSieveLane& l, uint32_t r) {
  uint32_t w = tab[(b.w[0] & 0xFFFFu) + r] & tab[(b.w[0] >> 16) + l.j[0]];
#pragma unroll
  for (int k = 1; k < kStrictRowWords; ++k)
    w &= tab[(b.w[k] & 0xFFFFu) + l.j[2 * k - 1]] & tab[(b.w[k] >> 16) + l.j[2 * k]];
  return w;
}

// ---------------------------------------------------------------------------
// The same sieves over consecutive c: a class 2-4 row (T(d) - c^6, c in the
// row's window) and a (6,1,7) twin row (T - c4^6 - c^6).  A word covers
// c = c0 .. c0+31; T - c^6 mod m depends on the row through A = X mod m and on
// c through c mod P, P the period of c^6 mod m: 2 for m = 8, 3 for 9, 32 for 64,
// 27 for 81, m for the others.  Reference moduli (LDS, 622 words) and strict
// moduli (global memory, 30595 words), tables [A][c0 mod P].
constexpr uint32_t kCSieveOff8 = 0;     // [A = X mod 8][c0 mod 2]
constexpr uint32_t kCSieveOff9 = 16;    // [A][c0 mod 3]
constexpr uint32_t kCSieveOff7 = 43;    // [A][c0 mod 7]
constexpr uint32_t kCSieveOff13 = 92;   // [A][c0 mod 13]
constexpr uint32_t kCSieveOff19 = 261;  // [A][c0 mod 19]
constexpr uint32_t kCSieveWords = 622;

struct CSieveTab {
  uint32_t w[kCSieveWords];
};

constexpr CSieveTab make_csieve() {
  CSieveTab s{};
  const uint32_t mods[5] = {8, 9, 7, 13, 19}, per[5] = {2, 3, 7, 13, 19};
  const uint32_t offs[5] = {kCSieveOff8, kCSieveOff9, kCSieveOff7, kCSieveOff13, kCSieveOff19};
  for (int t = 0; t < 5; ++t) {
    const uint32_t m = mods[t], P = per[t], g = sieve_sum_mask(m);
    for (uint32_t A = 0; A < m; ++A)
      for (uint32_t j0 = 0; j0 < P; ++j0) {
        uint32_t word = 0;
        for (uint32_t j = 0; j < 32; ++j)
          if ((g >> ((A + m - sieve_pow6((j0 + j) % P, m)) % m)) & 1u) word |= 1u << j;
        s.w[offs[t] + A * P + j0] = word;
      }
  }
  return s;
}

// the row's bases from xg = X mod 124488: ra = b8 | b9 << 16, rb = b7 | b13 << 16, rc = b19
FC5_HD void csieve_row_bases(uint32_t xg, uint32_t& ra, uint32_t& rb, uint32_t& rc) {
  ra = (kCSieveOff8 + (xg & 7u) * 2u) | (kCSieveOff9 + xg % 9u * 3u) << 16;
  rb = (kCSieveOff7 + xg % 7u * 7u) | (kCSieveOff13 + xg % 13u * 13u) << 16;
  rc = kCSieveOff19 + xg % 19u * 19u;
}

// the gate survivors among c = c0 .. c0+31 (before the window cut)
FC5_HD uint32_t csieve_word(const uint32_t* tab, uint32_t ra, uint32_t rb, uint32_t rc, uint32_t c0) {
  return tab[(ra & 0xFFFFu) + (c0 & 1u)] & tab[(ra >> 16) + c0 % 3u] & tab[(rb & 0xFFFFu) + c0 % 7u] &
         tab[(rb >> 16) + c0 % 13u] & tab[rc + c0 % 19u];
}

constexpr uint32_t kStrictCPer[kStrictMods] = {32, 27, 31, 37, 43, 49, 61, 67, 73, 79};

constexpr uint32_t strict_csieve_off(int i) {
  uint32_t o = 0;
  for (int k = 0; k < i; ++k) o += kStrictMod[k] * kStrictCPer[k];
  return o;
}
constexpr uint32_t kStrictCWords = strict_csieve_off(kStrictMods);
static_assert(kStrictCWords == 30595, "strict c-sieve size");

inline void fill_strict_csieve(uint32_t* w) {
  bool S[128];
  for (int t = 0; t < kStrictMods; ++t) {
    const uint32_t m = kStrictMod[t], P = kStrictCPer[t], off = strict_csieve_off(t);
    sieve_sum_set(m, S);
    for (uint32_t A = 0; A < m; ++A)
      for (uint32_t j0 = 0; j0 < P; ++j0) {
        uint32_t word = 0;
        for (uint32_t j = 0; j < 32; ++j)
          if (S[(A + m - sieve_pow6((j0 + j) % P, m)) % m]) word |= 1u << j;
        w[off + A * P + j0] = word;
      }
  }
}

// the row's strict c-sieve bases from its strict row words (X mod each modulus)
FC5_HD StrictSieveRow strict_csieve_bases(const StrictRow& xs) {
  const uint32_t xa = xs.w0 & 0xFFFFu, xb = xs.w1 & 0x3FFFFu, xc = xs.w1 >> 18;
  const uint32_t xm[kStrictMods] = {(xs.w0 >> 23) & 63u, (xs.w0 >> 16) & 127u, xa % 31u, xa % 37u, xa % 43u,
                                    xb % 49u, xb % 61u, xb % 67u, xc % 73u, xc % 79u};
  StrictSieveRow r;
#pragma unroll
  for (int k = 0; k < kStrictRowWords; ++k)
    r.w[k] = (strict_csieve_off(2 * k) + xm[2 * k] * kStrictCPer[2 * k]) |
             (strict_csieve_off(2 * k + 1) + xm[2 * k + 1] * kStrictCPer[2 * k + 1]) << 16;
  return r;
}

// the strict survivors among c = c0 .. c0+31
FC5_HD uint32_t strict_csieve_word(const uint32_t* tab, const StrictSieveRow& b, uint32_t c0) {
  uint32_t w = ~0u;
#pragma unroll
  for (int k = 0; k < kStrictRowWords; ++k)
    w &= tab[(b.w[k] & 0xFFFFu) + c0 % kStrictCPer[2 * k]] & tab[(b.w[k] >> 16) + c0 % kStrictCPer[2 * k + 1]];
  return w;
}

// the window cut: bits j with c0 + j <= h (c0 <= h)
FC5_HD uint32_t csieve_cut(uint32_t c0, uint32_t h) {
  const uint32_t jm = h - c0;
  return jm >= 31u ? 0xFFFFFFFFu : (2u << jm) - 1u;
}

}  // namespace fc5
