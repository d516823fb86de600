// Driver for the reference (6,1,7) solver's host code (solve_617_v1.cu built
// with HOST_ONLY), compiled against the reference's own file by
// oracle/Makefile into oracle/_ref/ref_617.  It only calls the reference's
// functions and prints their results, so tests can pin the oracle's twin
// restatement.
//
//   ref_617 cands B U_LO U_HI   -> one line per class-5 candidate of B
//                                  (gen_class_cands, solve_617_v1.cu:885-906):
//                                  "B u w base_hi base_lo lim fmax1 mod1 nres1 res..."
//   ref_617 tri                 -> triple-sum set sizes and bitmaps (solve_617_v1.cu:334-346)
#define main ref617_host_main
#include "solve_617_v1.cu"
#undef main

#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const RootTables rt;
  if (!std::strcmp(argv[1], "cands") && argc >= 5) {
    const long long B = std::atoll(argv[2]);
    std::vector<GpuCand> v;
    gen_class_cands(rt, B, 5, std::atof(argv[3]), std::atof(argv[4]), v);
    for (const GpuCand& C : v) {
      std::printf("%u %u %u %llu %llu %u %u %u %u", C.B, C.u, C.w, (unsigned long long)C.q_hi,
                  (unsigned long long)C.q_lo, C.lim, C.fmax1, C.mod1, C.nres1);
      for (u32 i = 0; i < C.nres1; ++i) std::printf(" %u", C.res[i]);
      std::printf("\n");
    }
    return 0;
  }
  if (!std::strcmp(argv[1], "tri")) {
    TriData td;
    build_tri_gate(td);
    std::printf("%d %d", td.n504, td.n247);
    for (int i = 0; i < 8; ++i) std::printf(" %llu", (unsigned long long)td.t504[i]);
    for (int i = 0; i < 4; ++i) std::printf(" %llu", (unsigned long long)td.t247[i]);
    std::printf("\n");
    return 0;
  }
  return 2;
}
