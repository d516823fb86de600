// Compiled by tests/test_native.py with g++: checks fc5_moddiv.hpp (the
// remainder used for xor cell indices on the device) against the % operator.
#include <cstdio>
#include <random>

#include "fc5_moddiv.hpp"

int main() {
  std::mt19937_64 rng(615);
  unsigned long long n = 0;
  auto check = [&](std::uint64_t h, std::uint64_t d) {
    const fc5::ModDiv m = fc5::make_moddiv(d);
    const std::uint64_t got = fc5::mod_fast(h, m), want = h % d;
    ++n;
    if (got != want) {
      std::printf("FAIL h=%llu d=%llu got=%llu want=%llu\n", (unsigned long long)h, (unsigned long long)d,
                  (unsigned long long)got, (unsigned long long)want);
      return false;
    }
    return true;
  };
  const std::uint64_t hs[] = {0, 1, 2, 3, ~0ULL, ~0ULL - 1, 1ULL << 63, (1ULL << 63) - 1, 1ULL << 32,
                              (1ULL << 32) - 1, (1ULL << 32) + 1, 0xFFFFFFFF00000000ULL};
  // divisors: every block size a table can have up to N = 120000, sampled, plus edges
  std::vector<std::uint64_t> ds = {1, 2, 3, 5, 7, 1000, 1209022, 65536, 65537, (1ULL << 31) - 1, 1ULL << 31,
                                   (1ULL << 31) + 1, (1ULL << 32) - 1, (1ULL << 32) - 2, 1ULL << 32,
                                   (1ULL << 32) + 7, 1ULL << 40, 0xFFFFFFFFFFFFFFFFULL};
  for (int i = 0; i < 2000; ++i) ds.push_back(1 + rng() % 3000000000ULL);
  for (int i = 0; i < 200; ++i) ds.push_back(1 + (rng() >> (rng() % 63)));
  for (std::uint64_t d : ds) {
    for (std::uint64_t h : hs)
      if (!check(h, d)) return 1;
    for (int i = 0; i < 2000; ++i) {
      const std::uint64_t h = rng();
      if (!check(h, d) || !check(h >> (i % 64), d) || !check((h / d) * d + (d - 1) * (i & 1), d)) return 1;
    }
  }
  std::printf("moddiv ok %llu cases\n", n);
  return 0;
}
