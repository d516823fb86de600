// fourcore_hunt_v3 -- Stage-1 unit emitter (host only).  Same flags and line
// formats as the reference program (fourcore_hunt_v3.cpp:251-451):
//   units (.buc):  "cls B u"
//   reduced T:     "cls B u free1 free2 T_lo T_hi"
//   fourcore_hunt_v3 [--selftest]
//   fourcore_hunt_v3 --density --lo A --hi B [--classes LIST] [--u-band lo,hi]
//   fourcore_hunt_v3 --lo A --hi B [--classes LIST] [--u-band lo,hi]
//                    [--emit-units FILE] [--expand --emit FILE] [--max-expand N]
// A FILE ending in .bbj is written in the binary .bbj layout (fourcore_job_bin.hpp):
// unit records for --emit-units, expanded jobs (free terms + exact T) for --emit.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fc5_host.hpp"

using namespace fc5;

static std::vector<int> parse_classes(const std::string& s) {
  std::vector<int> out;
  if (s.empty() || s == "all") return {1, 2, 3, 4, 5};
  for (char c : s)
    if (c >= '1' && c <= '5') out.push_back(c - '0');
  if (out.empty()) out = {1, 2, 3, 4, 5};
  return out;
}

static bool exact_T_cls1(u64 B, u64 u, u128& T) {  // (B^6-u^6)/42^6
  if (u >= B) return false;
  U256 num = U256::pow6(B);
  if (sub_in_place(num, U256::pow6(u)) || num.is_zero()) return false;
  if (divmod_small(num, kM42) != 0 || !num.fits128()) return false;
  T = num.lo128();
  return true;
}

static int selftest() {
  printf("=== fourcore_hunt_v3 selftest (host) ===\n");
  const i64 B = 100003;
  const size_t n1 = seeds_for_B(B, 1).size(), n2 = seeds_for_B(B, 2).size(), n3 = seeds_for_B(B, 3).size(),
               n4 = seeds_for_B(B, 4).size(), n5 = seeds_for_B(B, 5).size();
  printf("[1] seeds@B=%lld: cls1=%zu cls2=%zu cls3=%zu cls4=%zu cls5=%zu\n", (long long)B, n1, n2, n3, n4, n5);
  if (n1 != 144 || n2 != 24 || n3 != 36 || n4 != 6 || n5 != 6) {
    printf("FAIL seed cardinalities (want 144/24/36/6/6)\n");
    return 1;
  }
  const int b0 = U256::pow6(2353973).bits(), b1 = U256::pow6(2353974).bits();
  printf("[2] B^6 bits: 2353973 -> %d, 2353974 -> %d\n", b0, b1);
  if (b0 > 127 || b1 < 128) { printf("FAIL i128 boundary bits\n"); return 1; }
  // [5] cls5 double peel sample
  bool ok = false;
  for (i64 b = 100003; b < 102000 && !ok; ++b) {
    if (!admissible_B(b)) continue;
    for (i64 u : unit_candidates(b, 5, 0.05, 0.99)) {
      FreeSpec es, ds;
      free_de_cls5(b, u, es, ds);
      for (i64 e : es.res) {
        for (i64 d : ds.res) {
          u128 T;
          if (ok || !exact_T_cls5((u64)b, (u64)u, (u64)d, (u64)e, T) || T == 0) continue;
          ok = true;
          printf("[5] cls5 peel ok  B=%lld u=%lld d=%lld e=%lld\n", (long long)b, (long long)u, (long long)d,
                 (long long)e);
        }
      }
      if (ok) break;
    }
  }
  if (!ok) { printf("FAIL no cls5 peel sample\n"); return 1; }
  u64 cnt[6] = {0}, nB = 0;
  for (i64 b = 1000000; b <= 1000999; ++b) {
    if (!admissible_B(b)) continue;
    ++nB;
    for (int c = 1; c <= 5; ++c) cnt[c] += unit_candidates(b, c, 0.0, 1.0).size();
  }
  printf("[7] density B=[1000000,1000999] eligible_B=%llu units: c1=%llu c2=%llu c3=%llu c4=%llu c5=%llu\n",
         (unsigned long long)nB, (unsigned long long)cnt[1], (unsigned long long)cnt[2],
         (unsigned long long)cnt[3], (unsigned long long)cnt[4], (unsigned long long)cnt[5]);
  if (!(cnt[1] < cnt[3] && cnt[3] < cnt[2] && cnt[2] < cnt[4] && cnt[4] == cnt[5])) {
    printf("FAIL unexpected density ordering\n");
    return 1;
  }
  printf("SELFTEST PASS\n");
  return 0;
}

static void usage() {
  printf("usage: fourcore_hunt_v3 [--selftest]\n"
         "       fourcore_hunt_v3 --density --lo A --hi B [--classes LIST] [--u-band lo,hi]\n"
         "       fourcore_hunt_v3 --lo A --hi B [--classes LIST] [--u-band lo,hi]\n"
         "                        [--emit-units FILE] [--expand --emit FILE] [--max-expand N]\n");
}

int main(int argc, char** argv) {
  setvbuf(stdout, nullptr, _IOLBF, 0);
  setvbuf(stderr, nullptr, _IOLBF, 0);
  u64 lo = 0, hi = 0, max_expand = 0;
  double ulo = 0.0, uhi = 1.0;
  std::string classes_s = "all", emit_units, emit_t;
  bool do_self = false, do_density = false, do_expand = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string { return (i + 1 < argc) ? std::string(argv[++i]) : std::string(); };
    if (a == "--selftest") do_self = true;
    else if (a == "--density") do_density = true;
    else if (a == "--expand") do_expand = true;
    else if (a == "--lo") lo = std::strtoull(next().c_str(), nullptr, 10);
    else if (a == "--hi") hi = std::strtoull(next().c_str(), nullptr, 10);
    else if (a == "--classes") classes_s = next();
    else if (a == "--u-band") {
      const std::string s = next();
      if (std::sscanf(s.c_str(), "%lf,%lf", &ulo, &uhi) != 2) { fprintf(stderr, "bad --u-band\n"); return 1; }
    } else if (a == "--emit-units") emit_units = next();
    else if (a == "--emit") emit_t = next();
    else if (a == "--max-expand") max_expand = std::strtoull(next().c_str(), nullptr, 10);
    else if (a == "-h" || a == "--help") { usage(); return 0; }
    else { fprintf(stderr, "unknown %s\n", a.c_str()); usage(); return 1; }
  }
  if (do_self || argc == 1) return selftest();
  if (!lo || !hi || lo > hi) { usage(); return 1; }
  const auto classes = parse_classes(classes_s);
  if (do_density) {
    u64 nB = 0, units[6] = {0}, jobs[6] = {0};
    for (u64 B = lo; B <= hi; ++B) {
      if (!admissible_B((i64)B)) continue;
      ++nB;
      for (int cls : classes) {
        auto us = unit_candidates((i64)B, cls, ulo, uhi);
        units[cls] += us.size();
        for (i64 u : us) {
          if (cls == 1) jobs[1] += 1;
          else if (cls <= 4) {
            FreeSpec s = free_d_cls234((i64)B, u, cls);
            for (i64 r : s.res) for (i64 d = r; d <= s.fmax; d += s.mod) if (d) ++jobs[cls];
          } else {
            FreeSpec es, ds;
            free_de_cls5((i64)B, u, es, ds);
            u64 ne = 0, nd = 0;
            for (i64 r : es.res) for (i64 e = r; e <= es.fmax; e += es.mod) if (e) ++ne;
            for (i64 r : ds.res) for (i64 d = r; d <= ds.fmax; d += ds.mod) if (d) ++nd;
            jobs[5] += ne * nd;
          }
        }
      }
    }
    printf("density B=[%llu,%llu] u-band=[%.3f,%.3f) eligible_B=%llu\n", (unsigned long long)lo,
           (unsigned long long)hi, ulo, uhi, (unsigned long long)nB);
    for (int cls : classes)
      printf("  cls%d  units=%llu  expanded_jobs~%llu  (units/B=%.3f)\n", cls, (unsigned long long)units[cls],
             (unsigned long long)jobs[cls], nB ? (double)units[cls] / (double)nB : 0.0);
    return 0;
  }
  FILE* fu = nullptr;
  FILE* ft = nullptr;
  BbjWriter bu, bt;  // .bbj outputs (binary unit records / binary jobs)
  const bool u_bin = !emit_units.empty() && is_bbj_path(emit_units);
  const bool t_bin = do_expand && !emit_t.empty() && is_bbj_path(emit_t);
  if (!emit_units.empty()) {
    if (u_bin ? !bu.open(emit_units) : !(fu = fopen(emit_units.c_str(), "w"))) { perror(emit_units.c_str()); return 1; }
  }
  if (do_expand) {
    if (emit_t.empty()) { fprintf(stderr, "--expand requires --emit FILE\n"); return 1; }
    if (t_bin ? !bt.open(emit_t) : !(ft = fopen(emit_t.c_str(), "w"))) { perror(emit_t.c_str()); return 1; }
  } else if (!emit_t.empty()) {
    fprintf(stderr, "[warn] --emit without --expand ignored (use --emit-units or --expand)\n");
  }
  const bool want_units = fu || u_bin, want_jobs = ft || t_bin;
  if (!want_units && !want_jobs) {
    fprintf(stderr, "nothing to do: pass --emit-units and/or --expand --emit, or --density\n");
    usage();
    return 1;
  }
  u64 nB = 0, nUnits = 0, nJobs = 0, nFailT = 0;
  bool cap = false;
  fprintf(stderr, "[hunt_v3] B=[%llu,%llu] classes=", (unsigned long long)lo, (unsigned long long)hi);
  for (int c : classes) fprintf(stderr, "%d", c);
  fprintf(stderr, " u-band=[%.3f,%.3f) expand=%d\n", ulo, uhi, (int)do_expand);
  for (u64 B = lo; B <= hi && !cap; ++B) {
    if (!admissible_B((i64)B)) continue;
    ++nB;
    for (int cls : classes) {
      for (i64 u : unit_candidates((i64)B, cls, ulo, uhi)) {
        ++nUnits;
        if (fu) fprintf(fu, "%d %llu %lld\n", cls, (unsigned long long)B, (long long)u);
        if (u_bin) bu.add(cls, B, (u64)u, 0, 0, 0);
        if (!want_jobs) continue;
        auto emit = [&](u64 f1, u64 f2, u128 T) {
          if (max_expand && nJobs >= max_expand) { cap = true; return; }
          if (t_bin) bt.add(cls, B, (u64)u, f1, f2, T);
          else
            fprintf(ft, "%d %llu %lld %llu %llu %llu %llu\n", cls, (unsigned long long)B, (long long)u,
                    (unsigned long long)f1, (unsigned long long)f2, (unsigned long long)(u64)T,
                    (unsigned long long)(u64)(T >> 64));
          ++nJobs;
        };
        if (cls == 1) {
          u128 T;
          if (!exact_T_cls1(B, (u64)u, T)) { ++nFailT; continue; }
          emit(0, 0, T);
        } else if (cls <= 4) {
          const FreeSpec s = free_d_cls234((i64)B, u, cls);
          for (i64 r : s.res)
            for (i64 d = r; d <= s.fmax && !cap; d += s.mod) {
              if (!d) continue;
              u128 T;
              if (!exact_T_cls234(B, (u64)u, (u64)free_factor(cls), (u64)d, T)) { ++nFailT; continue; }
              if (T) emit((u64)d, 0, T);
            }
        } else {
          FreeSpec es, ds;
          free_de_cls5((i64)B, u, es, ds);
          for (i64 re : es.res)
            for (i64 e = re; e <= es.fmax && !cap; e += es.mod)
              for (i64 rd : ds.res)
                for (i64 d = rd; d <= ds.fmax && !cap; d += ds.mod) {
                  if (!d || !e) continue;
                  u128 T;
                  if (!exact_T_cls5(B, (u64)u, (u64)d, (u64)e, T)) { ++nFailT; continue; }
                  if (T) emit((u64)d, (u64)e, T);
                }
        }
        if (cap) break;
      }
      if (cap) break;
    }
    if ((B - lo) % 1000 == 0)
      fprintf(stderr, "[progress] B=%llu units=%llu jobs=%llu failT=%llu\n", (unsigned long long)B,
              (unsigned long long)nUnits, (unsigned long long)nJobs, (unsigned long long)nFailT);
  }
  fprintf(stderr, "[done] eligible_B=%llu units=%llu T_jobs=%llu failT=%llu%s\n", (unsigned long long)nB,
          (unsigned long long)nUnits, (unsigned long long)nJobs, (unsigned long long)nFailT,
          cap ? " (hit --max-expand)" : "");
  if (fu) fclose(fu);
  if (ft) fclose(ft);
  if ((u_bin && !bu.close()) || (t_bin && !bt.close())) {
    fprintf(stderr, "error writing the .bbj output\n");
    return 1;
  }
  return 0;
}
