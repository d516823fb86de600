// solve_617_v1 -- CLI front end over libfc5 for the (6,1,7) Branch-A class-5 twin.
//
// Same positional arguments, flags, exit codes and stdout/stderr line shapes as
// the reference program (solve_617_v1.cu:1686-2045) for the part this program
// provides, Branch A class 5:
//   solve_617_v1 <B_min> <B_max> 5 [u_lo] [u_hi] --branch-a-only [--chunk K]
//                [--device K] [--r R] [--hit-cap N] [--load-table F] [--save-table F]
//                [--no-gate] [--no-tri-gate] [--quiet] [--check-known [--known-file F]]
// Additions: --exact-c4 (the exact c4 bound 4 c4^6 >= T instead of the reference
// kernel's 3 c4^6 >= T), --ref-tiles (count and report as the reference launch
// does, whose unbounded 2048-wide c4 tiles search part of a window more than
// once past B ~ 515k), --gpu-build (build the table on the device instead of
// the host's byte-identical builder), --counters (one machine-readable line),
// --strict-gate (the strict residue gate on every key T - c4^6 - c3^6 before its
// query: the same solutions, fewer queries).
// Branch B, classes 1-4 of the (6,1,7) system, --xcheck and --bench are not
// provided here: asking for them is an error (exit 1).
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "fc5.h"

static constexpr int kXorNSoftMax = 120000;         // fourcore_xor_store.hpp:122
static constexpr long long kBHardMax = 2353973;     // solve_617_v1.cu:146 (B^6 < 2^127)

static void usage(const char* a0) {
  fprintf(stderr,
          "usage: %s <B_min> <B_max> [classes] [u_lo] [u_hi] [options]\n"
          "  (6,1,7): a1^6+..+a7^6 = B^6, primitive; this program runs Branch A class 5.\n"
          "  classes: \"5\"; u band as fractions of B\n"
          "  options: --branch-a-only --chunk K --device K --r R --hit-cap N\n"
          "           --save-table F --load-table F --quiet --no-gate --no-tri-gate\n"
          "           --check-known [--known-file F]  diff vs database after run\n"
          "           --exact-c4 --ref-tiles --gpu-build --counters --strict-gate\n",
          a0);
}

static int fail(fc5_ctx* ctx, const char* what) {
  fprintf(stderr, "[617] FATAL: %s: %s\n", what, ctx ? fc5_last_error(ctx) : "");
  if (ctx) fc5_ctx_destroy(ctx);
  return 1;
}

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  long long chunk = 8192;
  int device = 0, xor_r = 48;
  long long hit_cap = 1 << 20;
  std::string save_table, load_table, known_file = "617-solutions-clean.txt";
  bool quiet = false, no_gate = false, no_tri = false, check_known = false, exact_c4 = false;
  bool gpu_build = false, counters = false, branch_b = true, ref_tiles = false, strict = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--quiet") quiet = true;
    else if (a == "--no-gate") no_gate = true;
    else if (a == "--no-tri-gate") no_tri = true;
    else if (a == "--no-defer-gate-load") {}  // the batched kernel has no per-block gate load to defer
    else if (a == "--check-known") check_known = true;
    else if (a == "--branch-a-only") branch_b = false;
    else if (a == "--exact-c4") exact_c4 = true;
    else if (a == "--ref-tiles") ref_tiles = true;
    else if (a == "--gpu-build") gpu_build = true;
    else if (a == "--counters") counters = true;
    else if (a == "--strict-gate") strict = true;
    else if (a == "--known-file" && i + 1 < argc) known_file = argv[++i];
    else if (a == "--chunk" && i + 1 < argc) chunk = atoll(argv[++i]);
    else if (a == "--device" && i + 1 < argc) device = atoi(argv[++i]);
    else if (a == "--r" && i + 1 < argc) xor_r = atoi(argv[++i]);
    else if (a == "--hit-cap" && i + 1 < argc) hit_cap = atoll(argv[++i]);
    else if (a == "--save-table" && i + 1 < argc) save_table = argv[++i];
    else if (a == "--load-table" && i + 1 < argc) load_table = argv[++i];
    else if (a == "--branch-b-only" || a == "--xcheck" || a == "--selftest" || a == "--bench" ||
             a == "--branch-b-lim") {
      fprintf(stderr, "%s is not provided by this program (Branch A class 5 only)\n", a.c_str());
      return 1;
    } else if (!a.empty() && a[0] == '-') {
      fprintf(stderr, "unknown option %s\n", a.c_str());
      return 1;
    } else {
      pos.push_back(a);
    }
  }
  if (pos.size() < 2) {
    usage(argv[0]);
    return 1;
  }
  const long long B_min = atoll(pos[0].c_str()), B_max = atoll(pos[1].c_str());
  if (pos.size() >= 3 && pos[2] != "5") {
    fprintf(stderr, "classes \"%s\": this program runs class 5 only\n", pos[2].c_str());
    return 1;
  }
  if (branch_b) {
    fprintf(stderr, "Branch B (7 | B) is not provided by this program: pass --branch-a-only\n");
    return 1;
  }
  const double u_lo = pos.size() >= 4 ? atof(pos[3].c_str()) : 0.0;
  const double u_hi = pos.size() >= 5 ? atof(pos[4].c_str()) : 1.0;
  if (B_max > kBHardMax) {
    fprintf(stderr, "B_max capped at %lld (i128: B^6 < 2^127)\n", kBHardMax);
    return 1;
  }
  if (chunk < 1) chunk = 1;
  const int N = (int)(B_max / 42);
  if (N < 4) {
    fprintf(stderr, "B_max too small for Branch A\n");
    return 1;
  }

  std::set<long long> expect;
  if (check_known) {
    FILE* f = fopen(known_file.c_str(), "r");
    if (!f) fprintf(stderr, "[check-known] cannot open %s\n", known_file.c_str());
    size_t loaded = 0;
    char line[512];
    while (f && fgets(line, sizeof line, f)) {
      if (line[0] == '#' || line[0] == '\n') continue;
      long long B = 0;
      if (sscanf(line, "%lld", &B) != 1 || B <= 0) continue;
      ++loaded;
      if (B < B_min || B > B_max || B % 7 == 0) continue;
      // the terms after '=': a class-5 row has its odd, not-divisible-by-3 and
      // not-divisible-by-7 terms on three different terms
      std::vector<long long> t;
      const char* p = strchr(line, '=');
      while (p && *p && *p != '|') {
        char* end = nullptr;
        const long long x = strtoll(p + 1, &end, 10);
        if (end == p + 1) break;
        t.push_back(x);
        p = end;
        while (*p == ' ') ++p;
        if (*p != '+') break;
      }
      int n_odd = 0, n3 = 0, n7 = 0;
      long long o = 0, a3 = 0, a7 = 0;
      for (long long x : t) {
        if (x & 1) { ++n_odd; o = x; }
        if (x % 3) { ++n3; a3 = x; }
        if (x % 7) { ++n7; a7 = x; }
      }
      if (t.size() == 7 && n_odd == 1 && n3 == 1 && n7 == 1 && o != a3 && o != a7 && a3 != a7) expect.insert(B);
    }
    if (f) {
      fclose(f);
      fprintf(stderr, "[check-known] loaded %zu entries from %s\n", loaded, known_file.c_str());
    }
  }

  fc5_ctx* ctx = nullptr;
  if (fc5_ctx_create(device, &ctx) != FC5_OK) {
    fprintf(stderr, "no device %d\n", device);
    return 1;
  }
  fc5_set_option(ctx, "hit_cap", hit_cap);
  fc5_set_option(ctx, "use_gate", no_gate ? 0 : 1);
  fc5_set_option(ctx, "twin_tri_gate", no_tri ? 0 : 1);
  fc5_set_option(ctx, "twin_c4_div", exact_c4 ? 4 : 3);
  fc5_set_option(ctx, "twin_ref_tiles", ref_tiles ? 2048 : 0);
  fc5_set_option(ctx, "strict_gate", strict ? 1 : 0);
  bool loaded = false;
  if (!load_table.empty()) {
    const int rc = fc5_table_load(ctx, load_table.c_str(), N);
    if (rc == FC5_OK) loaded = true;
    else fprintf(stderr, "[617] --load-table %s: %s; building instead\n", load_table.c_str(), fc5_last_error(ctx));
  }
  if (!loaded) {
    if (N > kXorNSoftMax) {
      fprintf(stderr, "[617] N=%d > soft max %d (~B=%lld) - raise cap or shard\n", N, kXorNSoftMax,
              (long long)kXorNSoftMax * 42LL);
      fc5_ctx_destroy(ctx);
      return 1;
    }
    if (fc5_table_build(ctx, N, xor_r, nullptr, 0, gpu_build ? 1 : 0) != FC5_OK) return fail(ctx, "table build");
  }
  if (!save_table.empty() && fc5_table_save(ctx, save_table.c_str()) != FC5_OK) return fail(ctx, "save table");
  int fpr = 0;
  if (fc5_table_fpr_smoke(ctx, 200000, &fpr) != FC5_OK) return fail(ctx, "fpr smoke");
  if (fpr > 8) {  // solve_617_v1.cu:1828-1832
    fprintf(stderr, "[617] FATAL: xor FPR smoke too high (%d) - refuse to search\n", fpr);
    fc5_ctx_destroy(ctx);
    return 1;
  }
  fprintf(stderr, "solve_617_v1: B in [%lld, %lld], BranchA=on BranchB=off, classes 5, unit band [%.4f, %.4f)*B, "
          "chunk=%lld\n", B_min, B_max, u_lo, u_hi, chunk);

  long solutions = 0;
  long long cands = 0, exact = 0, fp_false = 0;
  unsigned long long calls = 0, gated = 0, probes = 0, tri_skip = 0, tri_tot = 0, maybe = 0;
  double kernel_ms = 0;
  std::set<std::array<long long, 7>> reported;
  std::set<long long> found_B;
  for (long long c0 = B_min; c0 <= B_max; c0 += chunk) {
    const long long c1 = std::min(B_max, c0 + chunk - 1);
    if (!quiet) {
      long long nA = 0;
      for (long long B = std::max(c0, 11LL); B <= c1; ++B) nA += (B & 1) && B % 3 && B % 7;
      fprintf(stderr, "[chunk] B %lld..%lld: Branch A (%lld values)\n", c0, c1, nA);
    }
    fc5_twin_result r;
    const int rc = fc5_twin_run_range(ctx, (uint64_t)std::max(c0, 11LL), (uint64_t)c1, u_lo, u_hi, &r);
    if (rc == FC5_E_OVERFLOW)
      fprintf(stderr, "!! hit buffer overflow at B<=%lld - rerun with larger --hit-cap\n", c1);
    else if (rc != FC5_OK)
      return fail(ctx, "twin run");
    cands += (long long)r.cands;
    calls += r.calls;
    gated += r.gated;
    probes += r.probes;
    tri_skip += r.c4_tri_skipped;
    tri_tot += r.c4_total;
    maybe += r.maybe;
    kernel_ms += r.kernel_ms;
    size_t nh = 0, ns = 0;
    fc5_twin_last_hits(ctx, nullptr, 0, &nh);
    std::vector<fc5_twin_hit> hits(nh);
    if (nh) fc5_twin_last_hits(ctx, hits.data(), nh, &nh);
    fc5_twin_last_solutions(ctx, nullptr, 0, &ns);
    std::vector<fc5_twin_solution> sols(ns);
    if (ns) fc5_twin_last_solutions(ctx, sols.data(), ns, &ns);
    for (const fc5_twin_hit& h : hits) {
      bool ok = false;
      for (const fc5_twin_solution& S : sols) ok |= S.d == h.d && S.c4 == h.c4 && S.c3 == h.c3;
      if (ok) ++exact;
      else ++fp_false;
    }
    for (const fc5_twin_solution& S : sols) {
      std::array<long long, 7> t;
      for (int i = 0; i < 7; ++i) t[(size_t)i] = (long long)S.terms[i];
      if (!reported.insert(t).second) continue;
      printf("SOLUTION cls=5 B=%llu a1=%lld a2=%lld a3=%lld a4=%lld a5=%lld a6=%lld a7=%lld (anchor=%llu) gcd=%lld %s\n",
             (unsigned long long)S.B, t[0], t[1], t[2], t[3], t[4], t[5], t[6], (unsigned long long)S.u,
             (long long)S.gcd, S.gcd == 1 ? "primitive" : "NON-PRIMITIVE");
      fflush(stdout);
      fprintf(stderr, "%llu\tSOLUTION\tcls=5\n", (unsigned long long)S.B);
      ++solutions;
      found_B.insert((long long)S.B);
    }
    if (!quiet)
      fprintf(stderr, "B %lld..%lld | cands c5=%lld | probes c5=%.2e | sol=%ld\n", c0, c1, cands, (double)probes,
              solutions);
  }
  fprintf(stderr, "---- summary ----\n");
  if (cands) {
    fprintf(stderr, "cls5: cands=%lld probes=%.4e kernel=%.1fs rate=%.2e probes/s exact=%lld fp-false=%lld "
            "gate-kill=%.2f%%", cands, (double)probes, kernel_ms / 1e3, probes / std::max(1e-9, kernel_ms / 1e3),
            exact, fp_false, 100.0 * (double)gated / std::max(1.0, (double)(gated + calls)));
    if (tri_tot)
      fprintf(stderr, " tri-skip=%.2f%% (%llu/%llu c4)", 100.0 * (double)tri_skip / (double)tri_tot, tri_skip,
              tri_tot);
    fprintf(stderr, "\n");
  }
  fprintf(stderr, "total solutions: %ld (unique B=%zu)\n", solutions, found_B.size());
  if (counters)
    printf("COUNTERS cands=%lld c4_total=%llu c4_tri_skipped=%llu calls=%llu gated=%llu probes=%llu maybe=%llu "
           "solutions=%ld\n", cands, tri_tot, tri_skip, calls, gated, probes, maybe, solutions);
  fc5_ctx_destroy(ctx);
  if (check_known) {  // solve_617_v1.cu:1278-1299; only class-5 rows can be found here
    int missing = 0, extra = 0;
    for (long long b : found_B)
      if (!expect.count(b)) {
        ++extra;
        fprintf(stderr, "[check-known] *** CANDIDATE NEW B=%lld (not in database - verify independently) ***\n", b);
      }
    for (long long b : expect)
      if (!found_B.count(b)) {
        ++missing;
        fprintf(stderr, "[check-known] MISSING B=%lld (regression - known solution not recovered!)\n", b);
      }
    fprintf(stderr, "[check-known] range [%lld,%lld]: expected=%zu class-5 found=%zu missing=%d new-candidates=%d %s\n",
            B_min, B_max, expect.size(), found_B.size(), missing, extra,
            missing == 0 ? (extra == 0 ? "PASS" : "PASS (with new candidates)") : "FAIL");
    return missing == 0 ? 0 : 1;
  }
  return 0;
}
