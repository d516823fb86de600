// Driver for the reference class-contract header (fourcore_classes_v3.hpp),
// compiled against the reference's own file by oracle/Makefile into
// oracle/_ref/ref_classes.  It only calls the reference's functions and
// prints their results, so tests can pin the oracle's Stage-1 restatement.
//
//   ref_classes units LO HI CLASSES            -> "cls B u" lines (huntv3:371-381)
//   ref_classes cls5 B u                       -> e-residues, d-residues, fmax
//   ref_classes cls234 B u CLS                 -> d-residues, mod, fmax
//   ref_classes seeds B                        -> seed counts for classes 1..5
#include "fourcore_classes_v3.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  fc3::RootTables rt;
  if (!std::strcmp(argv[1], "units") && argc >= 5) {
    const long long lo = std::atoll(argv[2]), hi = std::atoll(argv[3]);
    const char* cl = argv[4];
    for (long long B = lo; B <= hi; ++B) {
      if (!fc3::admissible_B(B)) continue;
      for (const char* c = cl; *c; ++c) {
        if (*c < '1' || *c > '5') continue;
        const int cls = *c - '0';
        for (long long u : fc3::unit_candidates(rt, B, cls, 0.0, 1.0))
          std::printf("%d %lld %lld\n", cls, B, u);
      }
    }
    return 0;
  }
  if (!std::strcmp(argv[1], "cls5") && argc >= 4) {
    const long long B = std::atoll(argv[2]), u = std::atoll(argv[3]);
    fc3::FreeTermSpec es, ds;
    fc3::free_de_cls5(rt, B, u, es, ds);
    std::printf("e %lld %lld", es.mod, es.fmax);
    for (long long r : es.residues) std::printf(" %lld", r);
    std::printf("\nd %lld %lld", ds.mod, ds.fmax);
    for (long long r : ds.residues) std::printf(" %lld", r);
    std::printf("\n");
    return 0;
  }
  if (!std::strcmp(argv[1], "cls234") && argc >= 5) {
    const long long B = std::atoll(argv[2]), u = std::atoll(argv[3]);
    const int cls = std::atoi(argv[4]);
    fc3::FreeTermSpec s = fc3::free_d_cls234(rt, B, u, cls);
    std::printf("d %lld %lld", s.mod, s.fmax);
    for (long long r : s.residues) std::printf(" %lld", r);
    std::printf("\n");
    return 0;
  }
  if (!std::strcmp(argv[1], "seeds") && argc >= 3) {
    const long long B = std::atoll(argv[2]);
    for (int cls = 1; cls <= 5; ++cls)
      std::printf("%zu%c", fc3::seeds_for_B(rt, B, cls).size(), cls == 5 ? '\n' : ' ');
    return 0;
  }
  return 2;
}
