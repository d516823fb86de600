// Driver for the reference's sorted pair index (QuadSumIndex, quad_sum.hpp:8-45),
// compiled against the reference's own header by oracle/Makefile into
// oracle/_ref/ref_quad.  It only builds the index and prints what the index
// holds, so tests can pin the exact sorted pair table.
//
//   ref_quad table N        -> "sum i j" per entry, in index order (sum in decimal)
//   ref_quad range N T...   -> "T lo hi" per target: equal_sum_range(T)
#include "quad_sum.hpp"

#include <cstdio>
#include <cstring>
#include <string>

static std::string dec(i128 v) {
  if (v == 0) return "0";
  std::string s;
  while (v > 0) {
    s.insert(s.begin(), (char)('0' + (int)(v % 10)));
    v /= 10;
  }
  return s;
}

static i128 parse(const char* s) {
  i128 v = 0;
  for (; *s; ++s) v = v * 10 + (*s - '0');
  return v;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  QuadSumIndex ix;
  ix.build(6, std::atoi(argv[2]));
  if (!std::strcmp(argv[1], "table")) {
    for (const PairEntry& e : ix.pairs) std::printf("%s %d %d\n", dec(e.sum).c_str(), e.i, e.j);
    return 0;
  }
  if (!std::strcmp(argv[1], "range")) {
    for (int k = 3; k < argc; ++k) {
      const auto [lo, hi] = ix.equal_sum_range(parse(argv[k]));
      std::printf("%s %zu %zu\n", argv[k], lo, hi);
    }
    return 0;
  }
  return 2;
}
