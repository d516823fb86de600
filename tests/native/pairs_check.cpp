// Compiled by tests/test_native.py with g++ against fc5_host.cpp: checks the
// exact pair search used by the hit verification against brute force.
#include <cstdio>
#include <random>

#include "fc5_host.hpp"

using namespace fc5;

static u128 p6(u64 x) { const u128 x2 = (u128)x * x; return x2 * x2 * x2; }

int main() {
  std::vector<std::pair<u32, u32>> got;
  exact_pairs(p6(3) + p6(4), 10, got);
  if (got.size() != 1 || got[0] != std::make_pair(3u, 4u)) { std::printf("FAIL 3,4\n"); return 1; }
  exact_pairs(2 * p6(5), 10, got);
  if (got.size() != 1 || got[0] != std::make_pair(5u, 5u)) { std::printf("FAIL 5,5\n"); return 1; }
  exact_pairs(p6(3) + p6(4), 3, got);  // b above the bound
  if (!got.empty()) { std::printf("FAIL bound\n"); return 1; }
  std::mt19937_64 rng(7);
  for (int t = 0; t < 3000; ++t) {
    const u64 a = 1 + rng() % 3000, b = a + rng() % 3000;
    const u128 T = p6(a) + p6(b) + (t % 3 == 0 ? 1 : 0);  // a third are not pair sums
    const u32 hi = (u32)(b + rng() % 4);
    exact_pairs(T, hi, got);
    std::vector<std::pair<u32, u32>> want;
    for (u64 x = 1; x <= hi; ++x)
      for (u64 y = x; y <= hi; ++y)
        if (p6(x) + p6(y) == T) want.emplace_back((u32)x, (u32)y);
    if (got != want) { std::printf("FAIL t=%d a=%llu b=%llu\n", t, (unsigned long long)a, (unsigned long long)b); return 1; }
    if (t > 200) break;  // the brute force is quadratic; the rest are covered by the sums above
  }
  // hit verification rejects a fingerprint that is not T mod 2^64
  std::vector<Solution> sols;
  if (verify_cls5_hit(2428, 100003, 17646, 2380, 9, 44, 12345, sols) != 0) { std::printf("FAIL fp\n"); return 1; }
  std::printf("pairs ok\n");
  return 0;
}
