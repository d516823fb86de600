// Host check of fc5_strict.hpp (the strict residue gate the kernels run):
// prints the ten residue masks for the Python side to compare with its
// brute-force sets, and checks the row/lane split T = X - L against the
// predicate on T's own residues for random T and lane values.
#include <cstdio>

#include "fc5_strict.hpp"

using namespace fc5;
using u128 = unsigned __int128;

static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t next() {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return s;
}

int main() {
  std::printf("mask31 %llx\nmask37 %llx\nmask43 %llx\nmask64 %llx\nmask81 %llx %llx\n",
              (unsigned long long)kSum31, (unsigned long long)kSum37, (unsigned long long)kSum43,
              (unsigned long long)kSum64, (unsigned long long)kSum81Lo, (unsigned long long)kSum81Hi);
  std::printf("mask49 %llx\nmask61 %llx\nmask67 %llx %llx\nmask73 %llx %llx\nmask79 %llx %llx\n",
              (unsigned long long)kSum49, (unsigned long long)kSum61, (unsigned long long)kSum67Lo,
              (unsigned long long)kSum67Hi, (unsigned long long)kSum73Lo, (unsigned long long)kSum73Hi,
              (unsigned long long)kSum79Lo, (unsigned long long)kSum79Hi);
  long bad = 0, passed = 0;
  for (int it = 0; it < 2000000; ++it) {
    const u128 T = ((u128)(next() >> 8) << 64) | next();  // below 2^120
    const uint32_t x = (uint32_t)(next() % (1u << 21));
    const bool want = strict_pass_t((uint32_t)(T % kStrictA), (uint32_t)(T % 81), (uint32_t)(T % 64),
                                    (uint32_t)(T % kStrictB), (uint32_t)(T % kStrictC));
    passed += want;
    // class 5: L = floor(x^6 / 64)
    const u128 x6 = (u128)x * x * x * x * x * x;
    u128 X = T + (x6 >> 6);
    StrictLane L = strict_lane_q64(x);
    if (strict_pass(strict_row((uint64_t)(X >> 64), (uint64_t)X), L) != want) ++bad;
    if (L.la != (uint32_t)((x6 >> 6) % kStrictA) || L.l81 != (uint32_t)((x6 >> 6) % 81) ||
        L.l64 != (uint32_t)((x6 >> 6) % 64) || L.lb != (uint32_t)((x6 >> 6) % kStrictB) ||
        L.lc != (uint32_t)((x6 >> 6) % kStrictC))
      ++bad;
    // classes 2-4: L = x^6
    X = T + x6;
    L = strict_lane_p6(x);
    if (strict_pass(strict_row((uint64_t)(X >> 64), (uint64_t)X), L) != want) ++bad;
    if (L.la != (uint32_t)(x6 % kStrictA) || L.l81 != (uint32_t)(x6 % 81) || L.l64 != (uint32_t)(x6 % 64) ||
        L.lb != (uint32_t)(x6 % kStrictB) || L.lc != (uint32_t)(x6 % kStrictC))
      ++bad;
  }
  // every sum of two sixth powers passes
  for (int it = 0; it < 1000000; ++it) {
    const uint64_t a = next() % (1u << 21), b = next() % (1u << 21);  // a^6 + b^6 < 2^127
    const u128 a3 = (u128)a * a * a, b3 = (u128)b * b * b;
    const u128 T = a3 * a3 + b3 * b3;
    if (!strict_pass_t((uint32_t)(T % kStrictA), (uint32_t)(T % 81), (uint32_t)(T % 64), (uint32_t)(T % kStrictB),
                       (uint32_t)(T % kStrictC)))
      ++bad;
  }
  std::printf("pass fraction %.4f\n", passed / 2e6);
  if (bad) {
    std::printf("strict mismatches %ld\n", bad);
    return 1;
  }
  std::printf("strict ok\n");
  return 0;
}
