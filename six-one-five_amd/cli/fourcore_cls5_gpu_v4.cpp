// fourcore_cls5_gpu_v4 -- CLI front end over libfc5 (include/fc5.h).
//
// Same flags, exit codes and stdout/stderr line shapes as the reference
// program (fourcore_cls5_gpu_v4.cu:596-709):
//   fourcore_cls5_gpu_v4 [--selftest-host]
//   fourcore_cls5_gpu_v4 --units FILE.buc [--r BITS] [--N K] [--quiet|--no-quiet]
//                        [--no-gate] [--stop-first] [--device K] [--load-table F]
//                        [--save-table F] [--allow-rebuild]
//   fourcore_cls5_gpu_v4 --check-table --N K --load-table F [--r BITS] [--save-table F]
// Additions: --mode plain|batched|chunked (default batched), --hit-cap N, --counters
// (one machine-readable counters line on stdout), --exact-table (check filter
// survivors against the exact sorted pair table of all a^6 + b^6, a <= b <= N,
// built on the device), --strict-gate (before each xor query also require T mod
// 64, 81, 31, 37, 43, 49, 61, 67, 73, 79 to be sums of two sixth powers: the same
// solutions, fewer queries, the filter false positives it rejects not reported).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fc5.h"
#include "fc5_host.hpp"

using namespace fc5;

static constexpr int kXorNSoftMax = 120000;  // fourcore_xor_store.hpp:122

static void usage() {
  printf("usage: fourcore_cls5_gpu_v4 [--selftest-host]\n"
         "   or: fourcore_cls5_gpu_v4 --units FILE.buc [--r BITS]\n"
         "                            [--N K] [--quiet|--no-quiet] [--no-gate]\n"
         "                            [--stop-first] [--device K]\n"
         "                            [--load-table F] [--save-table F]\n"
         "                            [--allow-rebuild] [--mode plain|batched|chunked]\n"
         "                            [--hit-cap N] [--counters] [--exact-table] [--strict-gate]\n"
         "   or: fourcore_cls5_gpu_v4 --check-table --N K --load-table F\n"
         "                            [--r BITS] [--save-table F] [--allow-rebuild]\n");
}

// host self-test: 192-bit powers, T-space unit form and funnel fingerprint
// against exact big-integer arithmetic (fourcore_cls5_gpu_v4.cu:186-263)
static int selftest_host() {
  printf("[selftest-host] cls5 peel (T-space form vs exact T)\n");
  for (u64 x : {1ull, 42ull, 100003ull, 2353975ull, 2752470ull}) {
    U256 p = U256::pow6(x), q = U256::of(x);
    for (int i = 0; i < 5; ++i) mul_small(q, x);
    if (cmp(p, q) != 0) { printf("FAIL pow6 x=%llu\n", (unsigned long long)x); return 1; }
  }
  auto check = [](u64 B, u64 u, u64 d, u64 e, const char* tag) -> int {
    u128 T;
    if (!exact_T_cls5(B, u, d, e, T)) { printf("FAIL exact T %s\n", tag); return 1; }
    Cls5Unit U;
    if (!pack_cls5(B, u, U) || !U.regular) { printf("FAIL pack %s\n", tag); return 1; }
    const u64 e3 = e * e * e, d3 = d * d * d;
    const u128 eq = ((u128)e3 * e3) / 729, dq = ((u128)d3 * d3) >> 6;
    if (U.Q - eq - dq != T) { printf("FAIL T-space %s\n", tag); return 1; }
    return 0;
  };
  if (check(100003, 17646, 9, 44, "pre-wall")) return 1;
  {
    Cls5Unit U;
    if (!pack_cls5(2353975, 85710, U)) { printf("FAIL pack past-wall unit\n"); return 1; }
    bool found = false;
    for (u32 ie = 0; ie < U.n1 && !found; ++ie)
      for (u32 id = 0; id < U.n2 && !found; ++id) {
        const u64 e = U.res1[ie] + 729ull * 7, d = U.res2[id] + 64ull * 7;
        u128 T;
        if (!exact_T_cls5(2353975, 85710, d, e, T)) continue;
        if (check(2353975, 85710, d, e, "past-wall")) return 1;
        found = true;
      }
    if (!found) { printf("FAIL no divisible past-wall (d,e)\n"); return 1; }
  }
  const int bits = U256::pow6(2752470).bits();
  printf("  pow6/T-space ok; past-wall peel ok; B=2752470 B^6 bits=%d\n", bits);
  if (bits > 192) { printf("FAIL >192 bits\n"); return 1; }
  printf("[selftest-host] PASS\n");
  return 0;
}

int main(int argc, char** argv) {
  setvbuf(stdout, nullptr, _IOLBF, 0);
  setvbuf(stderr, nullptr, _IOLBF, 0);
  std::string units_path, load_table, save_table, mode_s = "batched";
  int device = 0, N_override = 0, xor_r = 48;
  long long hit_cap = 1 << 20;
  bool do_host = false, quiet = true, use_gate = true, stop_first = false;
  bool allow_rebuild = false, check_table = false, counters = false, exact_table = false, strict = false;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    auto next = [&]() { return (i + 1 < argc) ? std::string(argv[++i]) : std::string(); };
    if (s == "--selftest-host") do_host = true;
    else if (s == "--units") units_path = next();
    else if (s == "--max-table-gb" || s == "--S") (void)next();  // ignored legacy flags
    else if (s == "--r") xor_r = std::atoi(next().c_str());
    else if (s == "--N") N_override = std::atoi(next().c_str());
    else if (s == "--quiet") quiet = true;
    else if (s == "--no-quiet") quiet = false;
    else if (s == "--no-gate") use_gate = false;
    else if (s == "--stop-first") stop_first = true;
    else if (s == "--device") device = std::atoi(next().c_str());
    else if (s == "--load-table") load_table = next();
    else if (s == "--save-table") save_table = next();
    else if (s == "--allow-rebuild") allow_rebuild = true;
    else if (s == "--check-table") check_table = true;
    else if (s == "--mode") mode_s = next();
    else if (s == "--hit-cap") hit_cap = std::atoll(next().c_str());
    else if (s == "--counters") counters = true;
    else if (s == "--exact-table") exact_table = true;
    else if (s == "--strict-gate") strict = true;
    else if (s == "-h" || s == "--help") { usage(); return 0; }
    else { fprintf(stderr, "unknown %s\n", s.c_str()); usage(); return 1; }
  }
  if (do_host || argc == 1) return selftest_host();
  if (xor_r < 8 || xor_r > 64) { fprintf(stderr, "--r must be 8..64\n"); return 1; }
  const int mode = mode_s == "plain" ? FC5_MODE_PLAIN : mode_s == "chunked" ? FC5_MODE_CHUNKED : FC5_MODE_BATCHED;
  if (mode_s != "plain" && mode_s != "batched" && mode_s != "chunked") {
    fprintf(stderr, "--mode must be plain, batched or chunked\n");
    return 1;
  }

  auto acquire = [&](int N, XorTable& t) -> bool {  // acquire_xor_filter (cls5v4:568-594)
    bool loaded = false;
    std::string msg;
    if (!load_table.empty()) {
      loaded = load_bin(load_table, N, t, msg) == kLoadOk;
      if (!loaded) {
        fprintf(stderr, "[xor] %s\n", msg.c_str());
        if (!allow_rebuild) {
          fprintf(stderr, "[cls5-gpu] FATAL: --load-table failed (see [xor] lines above). Fix N/path, or "
                          "pass --allow-rebuild to peel on the host (needs tens of GB RAM).\n");
          return false;
        }
        fprintf(stderr, "[cls5-gpu] --load-table failed; --allow-rebuild: peeling xor on host...\n");
      } else {
        fprintf(stderr, "[xor] loaded %s (N=%u, cells=%llu, r=%u, %.2f GB packed)\n", load_table.c_str(),
                t.hdr.N, (unsigned long long)t.hdr.m_cells, t.hdr.r, t.packed.size() / 1e9);
      }
    }
    if (!loaded) {
      fprintf(stderr, "[xor] building keys N=%d pairs=%.3e r=%d ...\n", N, (double)N * (N + 1) / 2.0, xor_r);
      if (!build_pairs_cpu(N, (u32)xor_r, {}, t, msg)) { fprintf(stderr, "[xor] %s\n", msg.c_str()); return false; }
    } else {
      xor_r = (int)t.hdr.r;
    }
    if (!save_table.empty()) {
      if (!save_bin(save_table, t, msg)) { fprintf(stderr, "[xor] %s\n", msg.c_str()); return false; }
      fprintf(stderr, "[xor] saved %s (%.2f GB packed, N=%u, r=%u)\n", save_table.c_str(),
              t.packed.size() / 1e9, t.hdr.N, t.hdr.r);
    }
    const int fpr = fpr_smoke(t);
    fprintf(stderr, "[xor] FPR smoke: %d/200000 random keys hit (expect ~0 at r=%u)\n", fpr, t.hdr.r);
    if (fpr < 0 || fpr > 16) {
      fprintf(stderr, "[cls5-gpu] FATAL: xor FPR smoke too high (%d) — refuse to search\n", fpr);
      return false;
    }
    return true;
  };

  if (check_table) {
    if (N_override <= 0) { fprintf(stderr, "--check-table needs --N\n"); return 1; }
    if (N_override > kXorNSoftMax) { fprintf(stderr, "N=%d exceeds soft xor cap %d\n", N_override, kXorNSoftMax); return 1; }
    XorTable t;
    if (!acquire(N_override, t)) return 1;
    fprintf(stderr, "[cls5-gpu] --check-table OK N=%d r=%d packed=%.3f GB\n", N_override, xor_r,
            t.packed.size() / 1e9);
    return 0;
  }
  if (units_path.empty()) { usage(); return 1; }

  std::vector<BucLine> lines;
  if (!read_buc(units_path, 1u << 5, lines) || lines.empty()) {
    fprintf(stderr, "no cls5 units in %s\n", units_path.c_str());
    return 1;
  }
  std::vector<fc5_unit> units;
  u32 need = 0;
  for (const BucLine& L : lines) {
    Cls5Unit U;
    if (!pack_cls5(L.B, L.u, U)) continue;
    need = std::max(need, U.lim);
    units.push_back({L.B, L.u});
  }
  fprintf(stderr, "[cls5-gpu] packed %zu/%zu units from %s\n", units.size(), lines.size(), units_path.c_str());
  if (units.empty()) return 1;
  const int N = N_override > 0 ? N_override : (int)need;
  if (N > kXorNSoftMax) { fprintf(stderr, "N=%d exceeds soft xor cap %d\n", N, kXorNSoftMax); return 1; }
  fprintf(stderr, "[cls5-gpu] N=%d r=%d\n", N, xor_r);
  XorTable t;
  if (!acquire(N, t)) return 1;

  fc5_ctx* ctx = nullptr;
  if (fc5_ctx_create(device, &ctx) != FC5_OK) {
    fprintf(stderr, "[cls5-gpu] FATAL: cannot open device %d\n", device);
    return 1;
  }
  fc5_set_option(ctx, "use_gate", use_gate ? 1 : 0);
  fc5_set_option(ctx, "strict_gate", strict ? 1 : 0);
  fc5_set_option(ctx, "stop_first", stop_first ? 1 : 0);
  fc5_set_option(ctx, "hit_cap", hit_cap);
  fc5_store_header h;
  static_assert(sizeof h == sizeof t.hdr, "header");
  std::memcpy(&h, &t.hdr, sizeof h);
  if (fc5_table_upload(ctx, &h, t.packed.data(), t.packed.size()) != FC5_OK) {
    fprintf(stderr, "[cls5-gpu] FATAL: %s\n", fc5_last_error(ctx));
    fc5_ctx_destroy(ctx);
    return 1;
  }
  XorTable().packed.swap(t.packed);  // host copy no longer needed
  if (exact_table) {
    const auto tx = std::chrono::steady_clock::now();
    uint64_t n_ex = 0;
    if (fc5_exact_build(ctx, N, nullptr, 0) != FC5_OK || fc5_exact_info(ctx, &n_ex, nullptr, nullptr) != FC5_OK) {
      fprintf(stderr, "[cls5-gpu] FATAL: %s\n", fc5_last_error(ctx));
      fc5_ctx_destroy(ctx);
      return 1;
    }
    fc5_set_option(ctx, "verify_table", 1);
    fprintf(stderr, "[cls5-gpu] exact sorted pair table: %llu entries (%.2f GB) built in %.1fs\n",
            (unsigned long long)n_ex, n_ex * 16 / 1e9,
            std::chrono::duration<double>(std::chrono::steady_clock::now() - tx).count());
  }
  const auto t0 = std::chrono::steady_clock::now();
  fc5_result r{};
  const int rc = fc5_run_units(ctx, units.data(), units.size(), mode, &r);
  if (rc != FC5_OK && rc != FC5_E_OVERFLOW) {
    fprintf(stderr, "[cls5-gpu] FATAL: %s\n", fc5_last_error(ctx));
    fc5_ctx_destroy(ctx);
    return 1;
  }
  if (rc == FC5_E_OVERFLOW) fprintf(stderr, "[cls5-gpu] WARNING: %s\n", fc5_last_error(ctx));
  std::vector<fc5_solution> sols(r.solutions);
  size_t ns = 0;
  fc5_last_solutions(ctx, sols.data(), sols.size(), &ns);
  for (const fc5_solution& S : sols)
    printf("SOLUTION cls=5 B=%llu u=%llu a1=%lld a2=%lld a3=%lld a4=%lld a5=%lld gcd=%lld %s\n",
           (unsigned long long)S.B, (unsigned long long)S.u, (long long)S.terms[0], (long long)S.terms[1],
           (long long)S.terms[2], (long long)S.terms[3], (long long)S.terms[4], (long long)S.gcd,
           S.gcd == 1 ? "primitive" : "NON-PRIMITIVE");
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double gate_pct = r.calls ? 100.0 * (double)r.gated / (double)r.calls : 0.0;
  (void)quiet;  // one run call: the progress line is printed once, at the end
  fprintf(stderr, "[cls5-gpu] units=%zu/%zu calls=%llu probes=%llu gate=%.1f%% sols=%llu %.1fs\n",
            units.size(), units.size(), (unsigned long long)r.calls, (unsigned long long)r.probes, gate_pct,
            (unsigned long long)r.solutions, sec);
  if (counters)
    printf("COUNTERS calls=%llu neg=%llu zero=%llu gated=%llu probes=%llu maybe=%llu side_maybe=%llu "
           "hits=%llu dropped=%llu solutions=%llu kernel_ms=%.3f strict_out=%llu\n",
           (unsigned long long)r.calls, (unsigned long long)r.neg, (unsigned long long)r.zero,
           (unsigned long long)r.gated, (unsigned long long)r.probes, (unsigned long long)r.maybe,
           (unsigned long long)r.side_maybe, (unsigned long long)r.n_hits, (unsigned long long)r.hits_dropped,
           (unsigned long long)r.solutions, r.kernel_ms, (unsigned long long)r.strict_out);
  printf("---- done cls5-gpu: solutions=%llu units=%zu calls=%llu probes=%llu time=%.1fs ----\n",
         (unsigned long long)r.solutions, units.size(), (unsigned long long)r.calls, (unsigned long long)r.probes,
         sec);
  fc5_ctx_destroy(ctx);
  return 0;
}
