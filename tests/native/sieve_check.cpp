// Host check of fc5_sieve.hpp (the bit sieve the class-5 sweep runs): for
// random rows X, residues r and grid positions k, the sieve bit of
// d = r + 64k equals the reference gate on the exact T = X - floor(d^6/64)
// (T mod 504 and T mod 247 sums of two sixth powers, sets by brute force),
// the strict sieve bit equals the strict gate (fc5_strict.hpp) on that T, the
// consecutive-c sieves (class 2-4 and twin rows, T = X - c^6) do the same for
// their words, and the cut masks keep exactly the grid points inside.
#include <cstdio>

#include "fc5_sieve.hpp"
#include "fc5_strict.hpp"

using namespace fc5;
using u128 = unsigned __int128;

static uint64_t s = 0x2545F4914F6CDD1DULL;
static uint64_t next() {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return s;
}

static bool g504[504], g247[247];

int main() {
  for (uint32_t a = 0; a < 504; ++a)
    for (uint32_t b = 0; b < 504; ++b) g504[(sieve_pow6(a, 504) + sieve_pow6(b, 504)) % 504] = true;
  for (uint32_t a = 0; a < 247; ++a)
    for (uint32_t b = 0; b < 247; ++b) g247[(sieve_pow6(a, 247) + sieve_pow6(b, 247)) % 247] = true;
  constexpr SieveTab tab = make_sieve();
  static uint32_t stab[kStrictWords];
  fill_strict_sieve(stab);
  long bad = 0, pass = 0, n = 0, spass = 0;
  for (int it = 0; it < 4000000; ++it) {
    const uint32_t r = (uint32_t)(next() & 63u);
    const uint32_t k = (uint32_t)(next() % 32768u);  // d < 2^21
    const uint32_t d = r + 64u * k;
    const uint64_t d3 = (uint64_t)d * d * d;
    const u128 d6 = (u128)d3 * d3;
    const uint32_t td = (uint32_t)(d6 & 63u);
    const u128 T = (((u128)(next() >> 12) << 64) | next()) + 1u;  // 1 <= T < 2^116 + 1
    const u128 X = T + (d6 >> 6);
    const bool want = g504[(uint32_t)(T % 504u)] && g247[(uint32_t)(T % 247u)];
    uint32_t ra, rb, rc;
    sieve_row_bases((uint32_t)(X % 124488u), td, ra, rb, rc);
    const uint32_t w = k >> 5, j = k & 31u, d0 = r + 2048u * w;
    const bool got = (sieve_word(tab.w, ra, rb, rc, r, d0) >> j) & 1u;
    if (got != want) ++bad;
    pass += want;
    ++n;
    const bool swant = strict_pass_t((uint32_t)(T % kStrictA), (uint32_t)(T % 81u), (uint32_t)(T % 64u),
                                     (uint32_t)(T % kStrictB), (uint32_t)(T % kStrictC));
    const StrictSieveRow sb = strict_sieve_bases(strict_row((uint64_t)(X >> 64), (uint64_t)X), td);
    const bool sgot = (strict_sieve_word(stab, sb, strict_sieve_lane(d0), r) >> j) & 1u;
    if (sgot != swant) ++bad;
    spass += swant;
  }
  // consecutive c: T = X - c^6, X random, c = c0 + j
  constexpr CSieveTab ctab = make_csieve();
  static uint32_t sctab[kStrictCWords];
  fill_strict_csieve(sctab);
  for (int it = 0; it < 2000000; ++it) {
    const uint32_t c = (uint32_t)(next() % (1u << 21));
    const uint64_t c3 = (uint64_t)c * c * c;
    const u128 T = (((u128)(next() >> 12) << 64) | next()) + 1u;
    const u128 X = T + (u128)c3 * c3;
    const uint32_t j = (uint32_t)(next() % 32u) % (c + 1u), c0 = c - j;  // c is bit j of the word at c0
    const bool want = g504[(uint32_t)(T % 504u)] && g247[(uint32_t)(T % 247u)];
    uint32_t ra, rb, rc;
    csieve_row_bases((uint32_t)(X % 124488u), ra, rb, rc);
    if ((((csieve_word(ctab.w, ra, rb, rc, c0) >> j) & 1u) != 0) != want) ++bad;
    const bool swant = strict_pass_t((uint32_t)(T % kStrictA), (uint32_t)(T % 81u), (uint32_t)(T % 64u),
                                     (uint32_t)(T % kStrictB), (uint32_t)(T % kStrictC));
    const StrictSieveRow sb = strict_csieve_bases(strict_row((uint64_t)(X >> 64), (uint64_t)X));
    if ((((strict_csieve_word(sctab, sb, c0) >> j) & 1u) != 0) != swant) ++bad;
  }
  for (uint32_t c0 = 0; c0 < 1000; c0 += 7)
    for (uint32_t h = c0; h < c0 + 80; ++h)
      for (uint32_t j = 0; j < 32; ++j)
        if ((((csieve_cut(c0, h) >> j) & 1u) != 0) != (c0 + j <= h)) ++bad;
  // the cut: bit j of sieve_cut(d0, dl) is set iff d0 + 64j <= dl
  for (int it = 0; it < 1000000; ++it) {
    const uint32_t d0 = (uint32_t)(next() % 3000000u), dl = (uint32_t)(next() % 3000000u);
    const uint32_t m = sieve_cut(d0, dl);
    for (uint32_t j = 0; j < 32; ++j)
      if ((((m >> j) & 1u) != 0) != ((uint64_t)d0 + 64u * j <= dl)) ++bad;
  }
  std::printf("gate pass fraction %.5f, strict %.5f over %ld points\n", (double)pass / n, (double)spass / n, n);
  if (bad) {
    std::printf("sieve mismatches %ld\n", bad);
    return 1;
  }
  std::printf("sieve ok\n");
  return 0;
}
