// Compiled by tests/test_native.py with g++ against fc5_host.cpp: the record
// a verified identity becomes (the SOLUTION line's fields), on fabricated term
// sets -- no (6,1,5) solution is known, so no real identity reaches it.  Checks
// the term order, the refusals (equal terms, a term not below B, a term below
// 1), the gcd over the terms and B, and the primitive / non-primitive split,
// for the class-5 shape (42a, 42b, 21d, 14e, u) and the class 2-4 shape
// (42a, 42b, 42c, f d, u).
#include <cstdio>

#include "fc5_host.hpp"

using namespace fc5;

static int fails = 0;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);        \
      ++fails;                                                \
    }                                                         \
  } while (0)

static bool same5(const i64 a[5], const i64 b[5]) {
  for (int i = 0; i < 5; ++i)
    if (a[i] != b[i]) return false;
  return true;
}

int main() {
  Solution s;
  {  // class 5: u = 12, d = 5, e = 7, a = 2, b = 3 -> terms 84, 126, 105, 98, 12
    const i64 t[5] = {42 * 2, 42 * 3, 21 * 5, 14 * 7, 12};
    CHECK(solution_record(1000003, 12, 5, 5, 7, 2, 3, 0, t, s));
    const i64 want[5] = {12, 84, 98, 105, 126};
    CHECK(same5(s.terms, want));
    CHECK(s.gcd == 1);  // gcd(12, 84) = 12, with 98 -> 2, with 105 -> 1: primitive
    CHECK(s.B == 1000003 && s.u == 12 && s.cls == 5 && s.d == 5 && s.e == 7 && s.a == 2 && s.b == 3 && s.c == 0);
  }
  {  // every term and B divisible by 5: NON-PRIMITIVE with gcd 5
    const i64 t[5] = {42 * 5, 42 * 10, 21 * 5, 14 * 15, 35};
    CHECK(!solution_record(5 * 200001, 35, 5, 5, 15, 5, 10, 0, t, s));  // 42*5 == 14*15: equal terms
    const i64 t2[5] = {42 * 5, 42 * 10, 21 * 5, 14 * 20, 35};
    CHECK(solution_record(5 * 200001, 35, 5, 5, 20, 5, 10, 0, t2, s));
    const i64 want[5] = {35, 105, 210, 280, 420};
    CHECK(same5(s.terms, want));
    CHECK(s.gcd == 5);  // gcd(35, 105, 210, 280, 420, 1000005): NON-PRIMITIVE
  }
  {  // gcd taken with B too: terms share 7, B does not
    const i64 t[5] = {42 * 1, 42 * 2, 21 * 3, 14 * 5, 7 * 11};
    CHECK(solution_record(1000003, 77, 5, 3, 5, 1, 2, 0, t, s));
    CHECK(s.gcd == 1);
    const i64 t2[5] = {42 * 1, 42 * 2, 21 * 3, 14 * 5, 7 * 11};
    CHECK(solution_record(7 * 142859, 77, 5, 3, 5, 1, 2, 0, t2, s));
    CHECK(s.gcd == 7);
  }
  {  // equal terms are refused (a = b, or 42a = 21d)
    const i64 t[5] = {42 * 3, 42 * 3, 21 * 5, 14 * 7, 12};
    CHECK(!solution_record(1000003, 12, 5, 5, 7, 3, 3, 0, t, s));
    const i64 t2[5] = {42 * 1, 42 * 3, 21 * 2, 14 * 7, 12};
    CHECK(!solution_record(1000003, 12, 5, 2, 7, 1, 3, 0, t2, s));
  }
  {  // the largest term must be below B
    const i64 t[5] = {42 * 2, 42 * 3, 21 * 5, 14 * 7, 12};
    CHECK(!solution_record(126, 12, 5, 5, 7, 2, 3, 0, t, s));
    CHECK(solution_record(127, 12, 5, 5, 7, 2, 3, 0, t, s));
  }
  {  // class 4 shape (f = 7): 42a, 42b, 42c, 7d, u; a term below 1 is refused
    const i64 t[5] = {42 * 4, 42 * 9, 42 * 6, 7 * 13, 18};
    CHECK(solution_record(2000003, 18, 4, 13, 0, 4, 9, 6, t, s));
    const i64 want[5] = {18, 91, 168, 252, 378};
    CHECK(same5(s.terms, want));
    CHECK(s.gcd == 1 && s.cls == 4 && s.c == 6);
    const i64 t0[5] = {42 * 4, 42 * 9, 42 * 6, 7 * 13, 0};
    CHECK(!solution_record(2000003, 0, 4, 13, 0, 4, 9, 6, t0, s));
  }
  if (fails) return 1;
  std::printf("solution ok\n");
  return 0;
}
