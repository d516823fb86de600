// fc5_host.hpp -- host-side core of the class-5 free-peel engine.
//
// Number theory of the Stage-1 contracts, exact big integers, the .buc unit
// files, the StoreHeader-v2 .bin pair tables, the xor-filter construction
// (CPU form, byte-identical to the reference builder), unit packing into the
// quotient ("T-space") form the device kernels use, and pair recovery /
// exact verification of filter hits.
//
// Reference interfaces mirrored here (same meaning, independent code):
//   fourcore_classes_v3.hpp:80-232   eligibility, seeds, unit filter, free terms
//   fourcore_hunt_v3.cpp:371-381     Stage-1 emission ("cls B u" lines)
//   fourcore_cls5_gpu_v4.cu:155-184  .buc reading, unit packing
//   spike/ribbon/store_header.hpp    64-byte header v2
//   spike/ribbon/xor_pack.hpp        r-bit cell packing (+8 byte pad)
//   spike/ribbon/xor_filter.hpp      3-hash xor filter build and query
//   fourcore_xor_store.hpp:23-181    pair keys, seed schedule, save/load, FPR smoke
//   fourcore_gmp.hpp:55-117          exact peel and identities (GMP replaced by U256)
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "fc5_mix.hpp"

namespace fc5 {

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i64 = std::int64_t;
using u128 = unsigned __int128;

constexpr i64 kM2 = 64, kM3 = 729, kM7 = 117649;
constexpr i64 kM14 = 7529536, kM21 = 85766121;
constexpr u64 kM42 = 5489031744ULL;  // 42^6
constexpr u64 kPhi = 0x9E3779B97F4A7C15ULL;
constexpr u32 kGateMod = 124488;      // 504 * 247: the gate depends on T mod this

// ---------------------------------------------------------------------------
// number theory (classes:37-232)
// ---------------------------------------------------------------------------
i64 pow6_mod(i64 x, i64 m);
bool admissible_B(i64 B);
// sorted sixth roots a (gcd(a,M)=1, 1<=a<M) of t modulo M in {64, 729, 7^6}
std::vector<i64> sixth_roots(i64 t, i64 M);
std::vector<i64> seeds_for_B(i64 B, int cls);
std::vector<i64> unit_candidates(i64 B, int cls, double lo_frac, double hi_frac);

struct FreeSpec {
  i64 mod = 0, fmax = 0;
  std::vector<i64> res;
};
void free_de_cls5(i64 B, i64 u, FreeSpec& e, FreeSpec& d);
FreeSpec free_d_cls234(i64 B, i64 u, int cls);
int free_factor(int cls);  // 14, 21, 7 for classes 2, 3, 4

// ---------------------------------------------------------------------------
// exact 256-bit unsigned arithmetic (replaces GMP on this path)
// ---------------------------------------------------------------------------
struct U256 {
  u64 l[4] = {0, 0, 0, 0};
  static U256 of(u128 v) { U256 r; r.l[0] = (u64)v; r.l[1] = (u64)(v >> 64); return r; }
  static U256 pow6(u64 x);  // x < 2^42
  bool is_zero() const { return !(l[0] | l[1] | l[2] | l[3]); }
  bool fits128() const { return !(l[2] | l[3]); }
  u128 lo128() const { return ((u128)l[1] << 64) | l[0]; }
  int bits() const;
};
bool sub_in_place(U256& a, const U256& b);  // returns borrow
void add_in_place(U256& a, const U256& b);
void mul_small(U256& a, u64 m);
u64 divmod_small(U256& a, u64 d);  // a /= d, returns a % d
int cmp(const U256& a, const U256& b);

// T = (B^6 - u^6 - (21d)^6 - (14e)^6) / 42^6 if positive, divisible and < 2^128
bool exact_T_cls5(u64 B, u64 u, u64 d, u64 e, u128& T);
// T = (B^6 - u^6 - (f*d)^6) / 42^6 (classes 2-4)
bool exact_T_cls234(u64 B, u64 u, u64 f, u64 d, u128& T);
bool verify_cls5(u64 B, u64 u, u64 d, u64 e, u64 a, u64 b);

// ---------------------------------------------------------------------------
// files
// ---------------------------------------------------------------------------
struct BucLine { int cls; u64 B, u; };
// Lines "cls B u"; non-matching lines are skipped; cls_mask bit c keeps class c.
// A path ending in .bbj is read as a binary file of unit records (below).
bool read_buc(const std::string& path, unsigned cls_mask, std::vector<BucLine>& out);

// Binary job/unit files in the .bbj layout (fourcore_job_bin.hpp:22-96): a
// 32-byte header {magic 'FCJB', version 2, n_jobs, rec_bytes 36, min_B, max_B}
// and 36-byte records {cls, pad[3], B, u, free1, free2, T_lo, T_hi}.  A unit
// record has free1 = free2 = 0 and T = 0; an expanded job carries its free
// terms and exact T.
bool is_bbj_path(const std::string& path);
class BbjWriter {
 public:
  bool open(const std::string& path);
  bool add(int cls, u64 B, u64 u, u64 free1, u64 free2, u128 T);
  bool close();  // patches the header with the count and B range
  ~BbjWriter();

 private:
  std::FILE* f_ = nullptr;
  u64 n_ = 0;
  u32 min_B_ = 0xffffffffu, max_B_ = 0;
};
// unit records of a .bbj file; records with free terms (expanded jobs) are
// refused with a message
bool read_units_bbj(const std::string& path, unsigned cls_mask, std::vector<BucLine>& out, std::string& msg);

#pragma pack(push, 1)
struct StoreHeader {
  u32 magic = 0x53353136u;
  u32 version = 2;
  u32 kind = 1;  // xor
  u32 r = 48;
  u32 w = 0;
  u32 N = 0;
  u64 n_keys = 0;
  u64 m_cells = 0;
  u64 mix_seed = 0;
  u32 shard_count = 1;
  u32 shard_index = 0;
  u32 shard_mode = 0;
  u32 reserved = 0;
};
#pragma pack(pop)
static_assert(sizeof(StoreHeader) == 64, "StoreHeader v2 is 64 bytes");

struct XorTable {
  StoreHeader hdr{};
  std::vector<u8> packed;  // r-bit cells + 8 pad bytes
};

enum LoadStatus { kLoadOk = 0, kLoadOpen = -1, kLoadBad = -2, kLoadVersion = -3,
                  kLoadNMismatch = -4, kLoadShort = -5 };
int load_bin(const std::string& path, int expected_N, XorTable& t, std::string& msg);
// the checks of load_bin (header, version, exact N, file long enough for the
// packed cells) without reading the cells: on kLoadOk *f is open just past the header
int open_bin(const std::string& path, int expected_N, StoreHeader& hdr, std::FILE*& f, std::string& msg);
bool save_bin(const std::string& path, const XorTable& t, std::string& msg);

// ---------------------------------------------------------------------------
// xor filter
// ---------------------------------------------------------------------------
inline u64 fp_mask(u32 r) { return r >= 64 ? ~0ULL : ((1ULL << r) - 1ULL); }
size_t packed_bytes(u64 m_cells, u32 r);
size_t capacity_for(size_t n);
u64 load_cell(const u8* p, u64 cell, u32 r);
void store_cell(u8* p, u64 cell, u32 r, u64 v);
void cell_indices(u64 h, u64 block, u64 out[3]);
bool xor_query(const XorTable& t, u64 key);
int fpr_smoke(const XorTable& t, int trials = 200000);
// keys: any order, duplicates allowed.  One construction attempt per seed in
// the schedule base + k*phi; returns false if all stall.
bool build_xor_cpu(std::vector<u64>& keys, u32 r, u64 base_seed, int seed_tries, XorTable& out);
// Pair table for 1<=a<=b<=N plus optional extra keys, with the reference's
// outer seed schedule (4 bases, one seed each).  Byte-identical to the
// reference builder when extras are empty.
bool build_pairs_cpu(int N, u32 r, const std::vector<u64>& extra, XorTable& out, std::string& msg);

// ---------------------------------------------------------------------------
// units in T-space
// ---------------------------------------------------------------------------
// For a regular class-5 unit every grid residual R = B^6-u^6-(21d)^6-(14e)^6
// is a multiple of 42^6 and T = R/42^6 = Q - floor(e^6/729) - floor(d^6/64),
// with Q = floor((B^6-u^6)/42^6) + c and c in {-1,0} fixed per unit (the
// remainders of e^6 mod 729 and d^6 mod 64 are constant over the residue
// classes).  Q < 2^128 for B < 2^25, so the device never needs 192 bits.
struct Cls5Unit {
  u64 B = 0, u = 0;
  U256 base;            // B^6 - u^6
  u128 Q = 0;           // T-space offset (regular units)
  u32 lim = 0, fmax1 = 0, fmax2 = 0;
  u32 n1 = 0, n2 = 0;
  u32 mod1 = 729, mod2 = 64;
  u32 te = 0, td = 0;   // e^6 mod 729 and d^6 mod 64 over the residue classes
  u32 res1[12] = {0}, res2[12] = {0};
  bool regular = false; // Q form valid (else per-point 192-bit path)
};
// false when the reference would not pack the unit (base<=0, >192 bits, no residues)
bool pack_cls5(u64 B, u64 u, Cls5Unit& U);

// calls of one unit in closed form: #valid e * #valid d
u64 cls5_calls(const Cls5Unit& U);

// Classes 2-4 (find3 leaf): T(d) = (B^6 - u^6 - (f d)^6) / 42^6 over the free
// term d = res + mod*k.  With f^6 * mod = 42^6 (f, mod) = (14, 729), (21, 64),
// (7, 46656) and d^6 mod `mod` constant over the residues, exactly
// T(d) = Q - floor(d^6 / mod), Q = floor((B^6-u^6)/42^6) + c, c in {-1, 0}.
struct Cls234Unit {
  int cls = 0;
  u64 B = 0, u = 0;
  U256 base;
  u128 Q = 0, inv = 1;  // inv: inverse of the odd part of mod, mod 2^128
  u32 f = 0, lim = 0, fmax = 0, mod = 0, dr = 0, shift = 0, nres = 0;
  u32 res[64] = {0};
  bool regular = false;
};
bool pack_cls234(int cls, u64 B, u64 u, Cls234Unit& U);
// Packs a unit list in parallel (OpenMP).  bu holds (B, u) pairs; cls may be
// null (all class 5).  slot[i] indexes u5 (class 5) or u3 (classes 2-4), -1
// for other classes; ok5/ok3 say whether the pack succeeded.
void pack_units(const u32* cls, const u64* bu, size_t n, std::vector<Cls5Unit>& u5, std::vector<char>& ok5,
                std::vector<Cls234Unit>& u3, std::vector<char>& ok3, std::vector<int>& slot);
u32 iroot6(u128 n);  // largest a with a^6 <= n
// the find3 window of a job (make_cand mode 3, fourcore_find_v4.cu:412-418):
// hi = min(iroot6(T), lim), lo = least c >= 1 with 3 c^6 >= T; empty if lo > hi
void find3_window(u128 T, u32 lim, u32& lo, u32& hi);
// jobs of one unit: (d, T) for every valid d with T > 0 (exact); failT counts
// the d whose residual is not positive or not divisible (fourcore_find_v4.cu:181-189)
void expand_cls234(const Cls234Unit& U, std::vector<std::pair<u32, u128>>& jobs, u64& failT);
bool verify_cls234(u64 B, u64 u, u64 f, u64 d, u64 a, u64 b, u64 c);

// ---------------------------------------------------------------------------
// pair recovery and exact verification (cls5v4:513-542, xstore:23-51)
// ---------------------------------------------------------------------------
struct Solution {
  u64 B, u;
  u32 cls, d, e, a, b, c;
  i64 terms[5];
  i64 gcd;
};
// The record of an identity that has been checked exactly (the SOLUTION line,
// fourcore_cls5_gpu_v4.cu:524-536, fourcore_find_v4.cu:427-457): the five terms
// sorted ascending; refused (false) when two terms are equal, the smallest is
// not positive or the largest is not below B; gcd = gcd of the terms and B
// (gcd 1: "primitive").
bool solution_record(u64 B, u64 u, u32 cls, u32 d, u32 e, u32 a, u32 b, u32 c, const i64 terms[5], Solution& out);
// Exact pairs a <= b <= hi with a^6 + b^6 == T, by two pointers over x^6.
// The reference recovers every pair with a^6+b^6 = fp (mod 2^64) and then keeps
// the exact ones (fourcore_cls5_gpu_v4.cu:517-522); fp = T mod 2^64 for every
// hit, so the exact pairs are the same set.
void exact_pairs(u128 T, u32 hi, std::vector<std::pair<u32, u32>>& out);
// full verify chain for one cls5 maybe-hit (table index cap N); appends solutions
int verify_cls5_hit(int N, u64 B, u64 u, u32 lim, u32 d, u32 e, u64 fp, std::vector<Solution>& out);
// find3 maybe-hit of class cls at (d, c3) (fourcore_find_v4.cu:535-551, :427-457)
int verify_find3_hit(int N, int cls, u64 B, u64 u, u32 d, u32 c, u64 fp, std::vector<Solution>& out);
// One maybe-hit to verify: cls 5 uses (d, e) with cap lim; classes 2-4 use (d, c3 = e).
struct VerifyJob {
  int cls;
  u64 B, u;
  u32 lim, d, e;
  u64 fp;
};
// Verifies jobs in parallel (OpenMP); solutions come out in job order.  With
// stop_first, output ends after the first job that yields a solution.
void verify_jobs(int N, const std::vector<VerifyJob>& jobs, bool stop_first, std::vector<Solution>& out);
// The verification in two halves around a pair search (done on the device by
// the library): the exact target T of the hit's pair part and the bound on
// its terms, then the identity and term checks of one found pair (a <= b).
struct PairTarget {
  u128 T;
  u32 hi;
  bool ok;  // false: the hit's fingerprint does not match an exact, positive T
};
PairTarget verify_target(int N, const VerifyJob& j);
int verify_finish(const VerifyJob& j, u32 a, u32 b, std::vector<Solution>& out);
// verify_target for every job (OpenMP)
void verify_targets(int N, const std::vector<VerifyJob>& jobs, std::vector<PairTarget>& out);
// verify_finish over every job's pairs in job order (stop_first as in verify_jobs)
void verify_finish_all(const std::vector<VerifyJob>& jobs, const std::vector<std::vector<std::pair<u32, u32>>>& pairs,
                       bool stop_first, std::vector<Solution>& out);

// ---------------------------------------------------------------------------
// (6,1,7) class-5 twin (solve_617_v1.cu)
// ---------------------------------------------------------------------------
// Candidate (B, u = 14u', w = 6v'): base = B^6 - u^6 - w^6 and the grid term
// 21d, d = res[i] + 64k <= fmax1.  Every residue has the same d^6 mod 64
// (td), so R = base - (21d)^6 is a multiple of 42^6 exactly when
// base = 21^6 td (mod 42^6), and then T(d) = R/42^6 = Q - floor(d^6/64) with
// Q = floor(base/42^6).
struct TwinCand {
  u64 B = 0, u = 0, w = 0;
  u128 base = 0, Q = 0;
  u32 lim = 0, fmax1 = 0, nres = 0, td = 0;
  u32 res[8] = {0};
};
struct TwinCandRaw {  // the generator's output, fields of GpuCand (solve_617_v1.cu:354-363)
  u64 B, u, w;
  u128 base;
  u32 lim, fmax1, mod1, nres1;
  u32 res[8];
};
// gen_class_cands cls 5 (solve_617_v1.cu:807-812,885-906)
void twin_cands(u64 B, double u_lo, double u_hi, std::vector<TwinCandRaw>& out);
// checks the T-space contract; false with a reason otherwise
bool pack_twin(const TwinCandRaw& in, TwinCand& out, std::string& why);
struct TwinSolution {
  u64 B, u, w;
  u32 d, c4, c3, a, b;
  i64 terms[7];
  i64 gcd;
};
// expand_xor_hits + verify_hit cls 5 (solve_617_v1.cu:1019-1110) for one hit;
// appends every exact, distinct, in-range 7-term identity (not deduplicated)
int verify_twin_hit(int N, const TwinCand& C, u32 d, u32 c4, u32 c3, u64 fp, std::vector<TwinSolution>& out);
PairTarget twin_target(int N, const TwinCand& C, u32 d, u32 c4, u32 c3, u64 fp);
int twin_finish(const TwinCand& C, u32 d, u32 c4, u32 c3, u32 a, u32 b, std::vector<TwinSolution>& out);
struct TwinHitIn {
  u32 cand, d, c4, c3;
  u64 fp;
};
// verify_twin_hit for every hit in parallel (OpenMP); per[i] gets hit i's solutions
void verify_twin_hits(int N, const std::vector<TwinCand>& cands, const std::vector<TwinHitIn>& hits,
                      std::vector<std::vector<TwinSolution>>& per);
// twin_target for every hit (OpenMP)
void twin_targets(int N, const std::vector<TwinCand>& cands, const std::vector<TwinHitIn>& hits,
                  std::vector<PairTarget>& out);

// splitmix64 (Vigna), the generator for planted keys
inline u64 splitmix64(u64& s) {
  u64 z = (s += kPhi);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace fc5
