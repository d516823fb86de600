// fc5_host.cpp -- host-side core (see fc5_host.hpp for the interface map).
#include "fc5_host.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace fc5 {

// ---------------------------------------------------------------------------
// number theory
// ---------------------------------------------------------------------------
i64 pow6_mod(i64 x, i64 m) {
  x %= m;
  if (x < 0) x += m;
  const i64 x2 = x * x % m;
  const i64 x3 = x2 * x % m;
  return x3 * x3 % m;
}

static i64 gcd_i64(i64 a, i64 b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) { const i64 t = a % b; a = b; b = t; }
  return a;
}

static i64 inv_mod(i64 a, i64 m) {  // extended Euclid, gcd(a,m)=1
  i64 t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr) {
    const i64 q = r / nr;
    i64 tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return t < 0 ? t + m : t;
}

bool admissible_B(i64 B) { return B > 1 && B % 2 != 0 && B % 3 != 0 && B % 7 != 0; }

// Sixth roots modulo M in {2^6, 3^6, 7^6}.  The unit group mod M is small
// and the roots of t are a0*z for one root a0 and every z with z^6 = 1, so we
// keep one root per residue and the roots of unity.
struct SixthRootTable {
  i64 M;
  std::vector<int32_t> one_root;  // one_root[t] = some unit a with a^6 = t, or -1
  std::vector<i64> unity;         // z with z^6 = 1, ascending
  explicit SixthRootTable(i64 m) : M(m), one_root((size_t)m, -1) {
    for (i64 a = 1; a < M; ++a) {
      if (gcd_i64(a, M) != 1) continue;
      const i64 t = pow6_mod(a, M);
      if (one_root[(size_t)t] < 0) one_root[(size_t)t] = (int32_t)a;
      if (t == 1) unity.push_back(a);
    }
  }
  std::vector<i64> roots(i64 t) const {
    std::vector<i64> out;
    t %= M;
    if (t < 0) t += M;
    const int32_t a0 = one_root[(size_t)t];
    if (a0 < 0) return out;
    out.reserve(unity.size());
    for (i64 z : unity) out.push_back((i64)a0 * z % M);
    std::sort(out.begin(), out.end());
    return out;
  }
};

static const SixthRootTable& table_for(i64 M) {
  static const SixthRootTable t2(kM2), t3(kM3), t7(kM7);
  if (M == kM2) return t2;
  if (M == kM3) return t3;
  if (M == kM7) return t7;
  throw std::invalid_argument("sixth_roots: modulus must be 64, 729 or 7^6");
}

std::vector<i64> sixth_roots(i64 t, i64 M) { return table_for(M).roots(t); }

static i64 crt_pair(i64 a, i64 ma, i64 b, i64 mb) {  // coprime moduli
  const i64 k = ((b - a) % mb + mb) % mb * inv_mod(ma % mb, mb) % mb;
  return a + ma * k;
}

std::vector<i64> seeds_for_B(i64 B, int cls) {
  const auto s7 = sixth_roots(pow6_mod(B, kM7), kM7);
  std::vector<i64> out;
  if (cls == 1) {
    const auto s2 = sixth_roots(pow6_mod(B, kM2), kM2);
    const auto s3 = sixth_roots(pow6_mod(B, kM3), kM3);
    for (i64 a : s2)
      for (i64 b : s3) {
        const i64 ab = crt_pair(a, kM2, b, kM3);
        for (i64 c : s7) out.push_back(crt_pair(ab, kM2 * kM3, c, kM7));
      }
  } else if (cls == 2) {
    for (i64 a : sixth_roots(pow6_mod(B, kM2), kM2))
      for (i64 c : s7) out.push_back(crt_pair(a, kM2, c, kM7));
  } else if (cls == 3) {
    for (i64 b : sixth_roots(pow6_mod(B, kM3), kM3))
      for (i64 c : s7) out.push_back(crt_pair(b, kM3, c, kM7));
  } else {
    out = s7;
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

static i64 class_modulus(int cls) {
  switch (cls) {
    case 1: return (i64)kM42;
    case 2: return kM14;
    case 3: return kM21;
    default: return kM7;
  }
}

static bool unit_ok(i64 u, int cls) {
  switch (cls) {
    case 1: return (u & 1) && u % 3 != 0 && u % 7 != 0;
    case 2: return (u & 1) && u % 3 == 0 && u % 9 != 0;
    case 3: return !(u & 1) && u % 3 != 0 && u % 7 != 0;
    case 4:
    case 5: return u % 6 == 0 && u % 7 != 0;
  }
  return false;
}

std::vector<i64> unit_candidates(i64 B, int cls, double lo_frac, double hi_frac) {
  const i64 M = class_modulus(cls);
  const i64 lo = std::max<i64>(1, (i64)(lo_frac * (double)B));
  const i64 hi = std::min<i64>(B - 1, (i64)(hi_frac * (double)B));
  std::vector<i64> out;
  if (lo > hi) return out;
  for (i64 s : seeds_for_B(B, cls)) {
    i64 u = ((s % M) + M) % M;
    if (u < lo) u += (lo - u + M - 1) / M * M;
    for (; u <= hi; u += M)
      if (unit_ok(u, cls)) out.push_back(u);
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

int free_factor(int cls) { return cls == 2 ? 14 : cls == 3 ? 21 : cls == 4 ? 7 : 0; }

void free_de_cls5(i64 B, i64 u, FreeSpec& e, FreeSpec& d) {
  const i64 te = pow6_mod(B, kM3) * inv_mod(pow6_mod(14, kM3), kM3) % kM3;
  e.res = sixth_roots(te, kM3);
  e.mod = kM3;
  e.fmax = (B - 1) / 14;
  const i64 rhs = ((pow6_mod(B, kM2) - pow6_mod(u, kM2)) % kM2 + kM2) % kM2;
  const i64 td = rhs * inv_mod(pow6_mod(21, kM2), kM2) % kM2;
  d.res = sixth_roots(td, kM2);
  d.mod = kM2;
  d.fmax = (B - 1) / 21;
}

FreeSpec free_d_cls234(i64 B, i64 u, int cls) {
  FreeSpec s;
  if (cls == 2) {
    s.res = sixth_roots(pow6_mod(B, kM3) * inv_mod(pow6_mod(14, kM3), kM3) % kM3, kM3);
    s.mod = kM3;
    s.fmax = (B - 1) / 14;
  } else if (cls == 3) {
    const i64 rhs = ((pow6_mod(B, kM2) - pow6_mod(u, kM2)) % kM2 + kM2) % kM2;
    s.res = sixth_roots(rhs * inv_mod(pow6_mod(21, kM2), kM2) % kM2, kM2);
    s.mod = kM2;
    s.fmax = (B - 1) / 21;
  } else if (cls == 4) {
    const auto a = sixth_roots(pow6_mod(B, kM2) * inv_mod(pow6_mod(7, kM2), kM2) % kM2, kM2);
    const auto b = sixth_roots(pow6_mod(B, kM3) * inv_mod(pow6_mod(7, kM3), kM3) % kM3, kM3);
    for (i64 x : a)
      for (i64 y : b) s.res.push_back(crt_pair(x, kM2, y, kM3));
    std::sort(s.res.begin(), s.res.end());
    s.mod = kM2 * kM3;
    s.fmax = (B - 1) / 7;
  }
  return s;
}

// ---------------------------------------------------------------------------
// U256
// ---------------------------------------------------------------------------
U256 U256::pow6(u64 x) {
  U256 r = U256::of(x);
  for (int i = 0; i < 5; ++i) mul_small(r, x);
  return r;
}

int U256::bits() const {
  for (int i = 3; i >= 0; --i)
    if (l[i]) return 64 * i + (64 - __builtin_clzll(l[i]));
  return 0;
}

bool sub_in_place(U256& a, const U256& b) {
  u64 borrow = 0;
  for (int i = 0; i < 4; ++i) {
    const u64 x = a.l[i], y = b.l[i];
    const u64 d = x - y;
    const u64 b1 = x < y;
    a.l[i] = d - borrow;
    borrow = b1 | (d < borrow);
  }
  return borrow != 0;
}

void add_in_place(U256& a, const U256& b) {
  u64 c = 0;
  for (int i = 0; i < 4; ++i) {
    const u128 s = (u128)a.l[i] + b.l[i] + c;
    a.l[i] = (u64)s;
    c = (u64)(s >> 64);
  }
}

void mul_small(U256& a, u64 m) {
  u64 c = 0;
  for (int i = 0; i < 4; ++i) {
    const u128 p = (u128)a.l[i] * m + c;
    a.l[i] = (u64)p;
    c = (u64)(p >> 64);
  }
}

u64 divmod_small(U256& a, u64 d) {
  u128 rem = 0;
  for (int i = 3; i >= 0; --i) {
    const u128 cur = (rem << 64) | a.l[i];
    a.l[i] = (u64)(cur / d);
    rem = cur % d;
  }
  return (u64)rem;
}

int cmp(const U256& a, const U256& b) {
  for (int i = 3; i >= 0; --i)
    if (a.l[i] != b.l[i]) return a.l[i] < b.l[i] ? -1 : 1;
  return 0;
}

static bool peel_T(U256 num, u128& T) {
  if (num.is_zero()) return false;
  if (divmod_small(num, kM42) != 0) return false;
  if (!num.fits128()) return false;
  T = num.lo128();
  return true;
}

bool exact_T_cls5(u64 B, u64 u, u64 d, u64 e, u128& T) {
  if (u >= B || d < 1 || e < 1) return false;
  U256 num = U256::pow6(B);
  if (sub_in_place(num, U256::pow6(u))) return false;
  if (sub_in_place(num, U256::pow6(21 * d))) return false;
  if (sub_in_place(num, U256::pow6(14 * e))) return false;
  return peel_T(num, T);
}

bool exact_T_cls234(u64 B, u64 u, u64 f, u64 d, u128& T) {
  if (u >= B || f < 1 || d < 1) return false;
  U256 num = U256::pow6(B);
  if (sub_in_place(num, U256::pow6(u))) return false;
  if (sub_in_place(num, U256::pow6(f * d))) return false;
  return peel_T(num, T);
}

bool verify_cls5(u64 B, u64 u, u64 d, u64 e, u64 a, u64 b) {
  U256 rhs = U256::pow6(u);
  add_in_place(rhs, U256::pow6(21 * d));
  add_in_place(rhs, U256::pow6(14 * e));
  U256 core = U256::pow6(a);
  add_in_place(core, U256::pow6(b));
  mul_small(core, kM42);
  add_in_place(rhs, core);
  return cmp(rhs, U256::pow6(B)) == 0;
}

// ---------------------------------------------------------------------------
// files
// ---------------------------------------------------------------------------
#pragma pack(push, 1)
struct BbjHeader {
  u32 magic = 0x4a424346u;  // 'FCJB'
  u32 version = 2;
  u64 n_jobs = 0;
  u32 rec_bytes = 36;
  u32 min_B = 0, max_B = 0;
  u32 pad = 0;
};
struct BbjRec {
  u8 cls = 0;
  u8 pad[3] = {0, 0, 0};
  u32 B = 0, u = 0, free1 = 0, free2 = 0;
  u64 T_lo = 0, T_hi = 0;
};
#pragma pack(pop)
static_assert(sizeof(BbjHeader) == 32 && sizeof(BbjRec) == 36, ".bbj layout");

bool is_bbj_path(const std::string& path) {
  if (path.size() < 4) return false;
  std::string e = path.substr(path.size() - 4);
  for (char& c : e) c = (char)std::tolower((unsigned char)c);
  return e == ".bbj";
}

bool BbjWriter::open(const std::string& path) {
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_) return false;
  BbjHeader h;
  return std::fwrite(&h, sizeof h, 1, f_) == 1;
}

bool BbjWriter::add(int cls, u64 B, u64 u, u64 free1, u64 free2, u128 T) {
  if (!f_ || B > 0xffffffffull || u > 0xffffffffull || free1 > 0xffffffffull || free2 > 0xffffffffull) return false;
  BbjRec r;
  r.cls = (u8)cls;
  r.B = (u32)B;
  r.u = (u32)u;
  r.free1 = (u32)free1;
  r.free2 = (u32)free2;
  r.T_lo = (u64)T;
  r.T_hi = (u64)(T >> 64);
  ++n_;
  min_B_ = std::min(min_B_, (u32)B);
  max_B_ = std::max(max_B_, (u32)B);
  return std::fwrite(&r, sizeof r, 1, f_) == 1;
}

bool BbjWriter::close() {
  if (!f_) return false;
  BbjHeader h;
  h.n_jobs = n_;
  h.min_B = n_ ? min_B_ : 0;
  h.max_B = max_B_;
  const bool ok = std::fseek(f_, 0, SEEK_SET) == 0 && std::fwrite(&h, sizeof h, 1, f_) == 1;
  const bool closed = std::fclose(f_) == 0;
  f_ = nullptr;
  return ok && closed;
}

BbjWriter::~BbjWriter() {
  if (f_) std::fclose(f_);
}

bool read_units_bbj(const std::string& path, unsigned cls_mask, std::vector<BucLine>& out, std::string& msg) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) { msg = "cannot open " + path; return false; }
  BbjHeader h;
  if (std::fread(&h, sizeof h, 1, f) != 1 || h.magic != 0x4a424346u || h.version != 2 || h.rec_bytes != 36) {
    std::fclose(f);
    msg = "bad .bbj header in " + path;
    return false;
  }
  BbjRec r;
  u64 n = 0;
  while (std::fread(&r, sizeof r, 1, f) == 1) {
    ++n;
    if (r.free1 || r.free2 || r.T_lo || r.T_hi) {
      std::fclose(f);
      msg = path + " holds expanded jobs (free terms set), not Stage-1 units";
      return false;
    }
    if (r.cls < 1 || r.cls > 5 || !(cls_mask & (1u << r.cls))) continue;
    out.push_back({(int)r.cls, (u64)r.B, (u64)r.u});
  }
  std::fclose(f);
  if (n != h.n_jobs) { msg = "short .bbj file " + path; return false; }
  return true;
}

bool read_buc(const std::string& path, unsigned cls_mask, std::vector<BucLine>& out) {
  if (is_bbj_path(path)) {
    std::string msg;
    if (read_units_bbj(path, cls_mask, out, msg)) return true;
    std::fprintf(stderr, "%s\n", msg.c_str());
    return false;
  }
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  char buf[256];
  while (std::fgets(buf, sizeof buf, f)) {
    int cls;
    unsigned long long B, u;
    if (std::sscanf(buf, "%d %llu %llu", &cls, &B, &u) != 3) continue;
    if (cls < 1 || cls > 5 || !(cls_mask & (1u << cls))) continue;
    out.push_back({cls, (u64)B, (u64)u});
  }
  std::fclose(f);
  return true;
}

int open_bin(const std::string& path, int expected_N, StoreHeader& hdr, std::FILE*& out, std::string& msg) {
  out = nullptr;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) { msg = "cannot open " + path + " for read"; return kLoadOpen; }
  StoreHeader h;
  const bool got = std::fread(&h, sizeof h, 1, f) == 1;
  if (!got || h.magic != 0x53353136u || h.version != 2 || h.shard_count < 1 ||
      h.shard_index >= h.shard_count || h.r == 0 || h.r > 64 || h.kind != 1) {
    std::fclose(f);
    if (got && h.magic == 0x53353136u && h.version != 2) {
      msg = "unsupported store version " + std::to_string(h.version) + " (need 2 packed)";
      return kLoadVersion;
    }
    msg = "bad file " + path;
    return kLoadBad;
  }
  if (expected_N > 0 && (int)h.N != expected_N) {
    std::fclose(f);
    msg = "N mismatch: file N=" + std::to_string(h.N) + ", need " + std::to_string(expected_N);
    return kLoadNMismatch;
  }
  // the header's cell count must match the bytes that follow before anything
  // is allocated for them (a damaged header would otherwise ask for any size)
  long here = std::ftell(f), end = -1;
  if (here >= 0 && std::fseek(f, 0, SEEK_END) == 0) end = std::ftell(f);
  if (end < 0 || std::fseek(f, here, SEEK_SET) != 0) {
    std::fclose(f);
    msg = "cannot size " + path;
    return kLoadShort;
  }
  if (h.m_cells > (~0ULL - 7) / 64 || packed_bytes(h.m_cells, h.r) > (u64)(end - here)) {
    std::fclose(f);
    msg = "short read " + path;
    return kLoadShort;
  }
  hdr = h;
  out = f;
  return kLoadOk;
}

int load_bin(const std::string& path, int expected_N, XorTable& t, std::string& msg) {
  t.packed.clear();
  FILE* f = nullptr;
  const int st = open_bin(path, expected_N, t.hdr, f, msg);
  if (st != kLoadOk) return st;
  const size_t want = packed_bytes(t.hdr.m_cells, t.hdr.r);
  t.packed.assign(want, 0);
  if (want && std::fread(t.packed.data(), 1, want, f) != want) {
    std::fclose(f);
    t.packed.clear();
    msg = "short read " + path;
    return kLoadShort;
  }
  std::fclose(f);
  return kLoadOk;
}

bool save_bin(const std::string& path, const XorTable& t, std::string& msg) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { msg = "cannot open " + path + " for write"; return false; }
  if (std::fwrite(&t.hdr, sizeof t.hdr, 1, f) != 1 ||
      (!t.packed.empty() && std::fwrite(t.packed.data(), 1, t.packed.size(), f) != t.packed.size())) {
    std::fclose(f);
    msg = "short write " + path;
    return false;
  }
  if (std::fclose(f) != 0) {  // buffered bytes that could not be written (a full disk)
    msg = "short write " + path;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// xor filter
// ---------------------------------------------------------------------------
size_t packed_bytes(u64 m_cells, u32 r) {
  if (m_cells == 0 || r == 0) return 0;
  return (size_t)((m_cells * (u64)r + 7) / 8 + 8);
}

size_t capacity_for(size_t n) {
  const double need = 32.0 + 1.23 * (double)n;
  size_t cap = (size_t)(need + 0.999);
  if (cap < 3) cap = 3;
  return (cap + 2) / 3 * 3;
}

u64 load_cell(const u8* p, u64 cell, u32 r) {
  const u64 bit = cell * (u64)r;
  const u64 byte = bit >> 3;
  const unsigned off = (unsigned)(bit & 7);
  u64 w;
  std::memcpy(&w, p + byte, 8);
  if (off + r <= 64) return (w >> off) & fp_mask(r);
  u64 w2;
  std::memcpy(&w2, p + byte + 8, 8);
  return ((w >> off) | (w2 << (64 - off))) & fp_mask(r);
}

void store_cell(u8* p, u64 cell, u32 r, u64 v) {
  const u64 mask = fp_mask(r);
  v &= mask;
  const u64 bit = cell * (u64)r;
  const u64 byte = bit >> 3;
  const unsigned off = (unsigned)(bit & 7);
  u64 w;
  std::memcpy(&w, p + byte, 8);
  if (off + r <= 64) {
    w = (w & ~(mask << off)) | (v << off);
    std::memcpy(p + byte, &w, 8);
    return;
  }
  const unsigned low = 64 - off;  // bits that land in the first word
  const u64 low_mask = (1ULL << low) - 1;
  w = (w & ~(low_mask << off)) | ((v & low_mask) << off);
  std::memcpy(p + byte, &w, 8);
  const unsigned high = r - low;
  const u64 high_mask = high >= 64 ? ~0ULL : ((1ULL << high) - 1);
  u64 w2;
  std::memcpy(&w2, p + byte + 8, 8);
  w2 = (w2 & ~high_mask) | ((v >> low) & high_mask);
  std::memcpy(p + byte + 8, &w2, 8);
}

void cell_indices(u64 h, u64 block, u64 out[3]) {
  out[0] = h % block;
  out[1] = mix64(h ^ kCellSalt1) % block + block;
  out[2] = mix64(h ^ kCellSalt2) % block + 2 * block;
}

bool xor_query(const XorTable& t, u64 key) {
  if (t.hdr.n_keys == 0 || t.packed.empty() || t.hdr.m_cells < 3) return false;
  const u64 block = t.hdr.m_cells / 3;
  const u64 h = mix64(key + t.hdr.mix_seed);
  u64 c[3];
  cell_indices(h, block, c);
  const u8* p = t.packed.data();
  return (load_cell(p, c[0], t.hdr.r) ^ load_cell(p, c[1], t.hdr.r) ^ load_cell(p, c[2], t.hdr.r)) ==
         (h & fp_mask(t.hdr.r));
}

int fpr_smoke(const XorTable& t, int trials) {
  if (t.packed.empty() || t.hdr.m_cells < 3 || t.hdr.r == 0) return -1;
  int hits = 0;
  u64 x = 0x123456789abcdef0ULL;
  for (int i = 0; i < trials; ++i) {
    x = x * kPhi + 1;
    hits += xor_query(t, x) ? 1 : 0;
  }
  return hits;
}

// Parallel LSD radix sort (11-bit digits, 6 passes) followed by unique.
static void sort_unique(std::vector<u64>& v) {
  const size_t n = v.size();
  if (n < (1u << 16)) {
    std::sort(v.begin(), v.end());
  } else {
    std::vector<u64> tmp(n);
    u64* src = v.data();
    u64* dst = tmp.data();
    constexpr int kBits = 11, kR = 1 << kBits;
    int nt = 1;
#ifdef _OPENMP
    nt = std::min(omp_get_max_threads(), 64);
#endif
    std::vector<size_t> hist((size_t)nt * kR);
    for (int pass = 0; pass < 6; ++pass) {
      const int sh = pass * kBits;
      std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
      {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const size_t lo = n * (size_t)t / (size_t)nt, hi = n * (size_t)(t + 1) / (size_t)nt;
        size_t* h = &hist[(size_t)t * kR];
        for (size_t i = lo; i < hi; ++i) h[(src[i] >> sh) & (kR - 1)]++;
#pragma omp barrier
#pragma omp single
        {
          size_t s = 0;
          for (int b = 0; b < kR; ++b)
            for (int q = 0; q < nt; ++q) {
              const size_t c = hist[(size_t)q * kR + b];
              hist[(size_t)q * kR + b] = s;
              s += c;
            }
        }
        for (size_t i = lo; i < hi; ++i) dst[h[(src[i] >> sh) & (kR - 1)]++] = src[i];
      }
      std::swap(src, dst);
    }
    // 6 passes: data is back in v
  }
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// One construction attempt (xor_filter.hpp:256-431 semantics).  Incidence is
// order-free (xor, add), so it runs in parallel; the FIFO peel and the
// reverse assignment are sequential because they fix the table bytes.
static int construct_once(const std::vector<u64>& keys, u32 r, u64 seed, XorTable& out) {
  const size_t n = keys.size();
  out.hdr = StoreHeader{};
  out.hdr.r = r;
  out.hdr.mix_seed = seed;
  if (n == 0) { out.packed.clear(); return 1; }
  const size_t m = capacity_for(n);
  const u64 block = m / 3;
  std::vector<u64> setxor(m, 0);
  std::vector<u32> counts(m, 0);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    const u64 h = mix64(keys[i] + seed);
    u64 c[3];
    cell_indices(h, block, c);
    for (int t = 0; t < 3; ++t) {
      __atomic_fetch_xor(&setxor[c[t]], h, __ATOMIC_RELAXED);
      __atomic_fetch_add(&counts[c[t]], 1u, __ATOMIC_RELAXED);
    }
  }
  for (size_t i = 0; i < m; ++i)
    if (counts[i] > 0xffffu) return -1;  // the reference's u16 incidence overflow
  std::vector<std::pair<u64, u64>> stack;
  stack.reserve(n);
  std::vector<u64> queue;
  queue.reserve(1 << 20);
  for (size_t i = 0; i < m; ++i)
    if (counts[i] == 1) queue.push_back(i);
  for (size_t q = 0; q < queue.size(); ++q) {
    const u64 i = queue[q];
    if (counts[i] != 1) continue;
    const u64 h = setxor[i];
    u64 c[3];
    cell_indices(h, block, c);
    stack.emplace_back(h, i);
    for (int t = 0; t < 3; ++t) {
      const u64 j = c[t];
      setxor[j] ^= h;
      if (counts[j] > 0 && --counts[j] == 1) queue.push_back(j);
    }
  }
  if (stack.size() != n) return 0;
  std::vector<u64>().swap(setxor);
  std::vector<u32>().swap(counts);
  std::vector<u64>().swap(queue);
  const u64 mask = fp_mask(r);
  out.packed.assign(packed_bytes(m, r), 0);
  u8* p = out.packed.data();
  for (size_t k = stack.size(); k-- > 0;) {
    const u64 h = stack[k].first, at = stack[k].second;
    u64 c[3];
    cell_indices(h, block, c);
    u64 v = h & mask;
    for (int t = 0; t < 3; ++t)
      if (c[t] != at) v ^= load_cell(p, c[t], r);
    store_cell(p, at, r, v);
  }
  out.hdr.n_keys = n;
  out.hdr.m_cells = m;
  return 1;
}

bool build_xor_cpu(std::vector<u64>& keys, u32 r, u64 base_seed, int seed_tries, XorTable& out) {
  sort_unique(keys);
  for (int t = 0; t < seed_tries; ++t) {
    const int rc = construct_once(keys, r, base_seed + (u64)t * kPhi, out);
    if (rc == 1) return true;
    if (rc < 0) return false;
  }
  return false;
}

bool build_pairs_cpu(int N, u32 r, const std::vector<u64>& extra, XorTable& out, std::string& msg) {
  if (N < 1 || r < 1 || r > 64) { msg = "bad N or r"; return false; }
  std::vector<u64> pw6((size_t)N + 1);
  for (int x = 1; x <= N; ++x) { const u64 x2 = (u64)x * (u64)x; pw6[(size_t)x] = x2 * x2 * x2; }
  const size_t np = (size_t)N * (size_t)(N + 1) / 2;
  for (int outer = 0; outer < 4; ++outer) {
    std::vector<u64> keys(np + extra.size());
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 1; i <= N; ++i) {
      // row i starts after rows 1..i-1, each of length N-j+1
      size_t k = (size_t)(i - 1) * (size_t)N - (size_t)(i - 1) * (size_t)(i - 2) / 2;
      for (int j = i; j <= N; ++j) keys[k++] = pw6[(size_t)i] + pw6[(size_t)j];
    }
    std::copy(extra.begin(), extra.end(), keys.begin() + (ptrdiff_t)np);
    const u64 base = 0x615615615615615ULL + (u64)outer * kPhi;
    if (build_xor_cpu(keys, r, base, 1, out)) {
      out.hdr.N = (u32)N;
      return true;
    }
  }
  msg = "xor construction stalled at every seed";
  return false;
}

// ---------------------------------------------------------------------------
// units
// ---------------------------------------------------------------------------
bool pack_cls5(u64 B, u64 u, Cls5Unit& U) {
  U = Cls5Unit{};
  U.B = B;
  U.u = u;
  U.lim = (u32)((B - 1) / 42);
  if (u >= B || B >= (1ULL << 32)) return false;
  U.base = U256::pow6(B);
  if (sub_in_place(U.base, U256::pow6(u)) || U.base.is_zero()) return false;
  FreeSpec es, ds;
  free_de_cls5((i64)B, (i64)u, es, ds);
  U.mod1 = (u32)es.mod;
  U.mod2 = (u32)ds.mod;
  U.fmax1 = (u32)es.fmax;
  U.fmax2 = (u32)ds.fmax;
  U.n1 = (u32)std::min<size_t>(es.res.size(), 12);
  U.n2 = (u32)std::min<size_t>(ds.res.size(), 12);
  for (u32 i = 0; i < U.n1; ++i) U.res1[i] = (u32)es.res[i];
  for (u32 i = 0; i < U.n2; ++i) U.res2[i] = (u32)ds.res[i];
  if (U.n1 == 0 || U.n2 == 0) return false;
  // T-space form: every e residue has the same e^6 mod 729 and every d residue
  // the same d^6 mod 64; then R = 42^6 (bq - eq - dq) + (br - 14^6 te - 21^6 td).
  U.te = (u32)pow6_mod((i64)U.res1[0], kM3);
  U.td = (u32)pow6_mod((i64)U.res2[0], kM2);
  bool uniform = true;
  for (u32 i = 0; i < U.n1; ++i) uniform &= pow6_mod((i64)U.res1[i], kM3) == (i64)U.te;
  for (u32 i = 0; i < U.n2; ++i) uniform &= pow6_mod((i64)U.res2[i], kM2) == (i64)U.td;
  U256 q = U.base;
  const u64 br = divmod_small(q, kM42);
  // 14^6 = kM14 and 21^6 = kM21; the remainder sum is a multiple of 42^6 in (-2*42^6, 42^6)
  const __int128 rem = (__int128)br - (__int128)kM14 * U.te - (__int128)kM21 * U.td;
  const bool divisible = rem == 0 || rem == -(__int128)kM42;
  const bool small = q.fits128() && U.fmax1 < (1u << 21) && U.fmax2 < (1u << 21) && U.n1 == 6 &&
                     U.n2 == 4 && U.mod1 == 729 && U.mod2 == 64;
  U.regular = uniform && divisible && small;
  if (U.regular) U.Q = q.lo128() - (rem == 0 ? 0 : 1);
  return true;
}

u64 cls5_calls(const Cls5Unit& U) {
  auto count = [](const u32* res, u32 n, u32 mod, u32 fmax) {
    u64 c = 0;
    for (u32 i = 0; i < n; ++i)
      for (u64 x = res[i]; x <= fmax; x += mod)
        if (x) ++c;
    return c;
  };
  return count(U.res1, U.n1, U.mod1, U.fmax1) * count(U.res2, U.n2, U.mod2, U.fmax2);
}

static u128 inv_odd_2_128(u128 a) {  // a odd; Newton iteration doubles the correct bits
  u128 x = a;
  for (int i = 0; i < 7; ++i) x *= (u128)2 - a * x;
  return x;
}

bool pack_cls234(int cls, u64 B, u64 u, Cls234Unit& U) {
  U = Cls234Unit{};
  U.cls = cls;
  U.B = B;
  U.u = u;
  U.lim = (u32)((B - 1) / 42);
  if (cls < 2 || cls > 4 || B >= (1ULL << 32)) return false;
  const FreeSpec s = free_d_cls234((i64)B, (i64)u, cls);
  if (s.mod <= 0 || s.fmax < 1 || s.res.empty() || s.res.size() > 64) return false;
  U.f = (u32)free_factor(cls);
  U.mod = (u32)s.mod;
  U.fmax = (u32)s.fmax;
  U.nres = (u32)s.res.size();
  for (u32 i = 0; i < U.nres; ++i) U.res[i] = (u32)s.res[i];
  // B^6 - u^6 <= 0: the reference still walks the free terms and counts every
  // one as failT (fourcore_find_v4.cu:179-185); the per-term path does the same
  U.base = U256::pow6(B);
  if (u >= B || sub_in_place(U.base, U256::pow6(u)) || U.base.is_zero()) {
    U.base = U256{};
    return true;
  }
  U.shift = (u32)__builtin_ctz(U.mod);
  U.inv = inv_odd_2_128((u128)(U.mod >> U.shift));
  U.dr = (u32)pow6_mod((i64)U.res[0], (i64)U.mod);
  bool uniform = true;
  for (u32 i = 0; i < U.nres; ++i) uniform &= pow6_mod((i64)U.res[i], (i64)U.mod) == (i64)U.dr;
  U256 q = U.base;
  const u64 br = divmod_small(q, kM42);
  const u64 f6 = (u64)U.f * U.f * U.f;
  const __int128 rem = (__int128)br - (__int128)(f6 * f6) * U.dr;
  const bool divisible = rem == 0 || rem == -(__int128)kM42;
  U.regular = uniform && divisible && q.fits128() && U.fmax < (1u << 21) && U.nres <= 24;
  if (U.regular) U.Q = q.lo128() - (rem == 0 ? 0 : 1);
  return true;
}

u32 iroot6(u128 n) {
  const double x = (double)(u64)(n >> 64) * 18446744073709551616.0 + (double)(u64)n;
  if (x < 1.0) return 0;
  u64 a = (u64)std::cbrt(std::sqrt(x));
  auto p6 = [](u64 v) { const u128 v2 = (u128)v * v; return v2 * v2 * v2; };
  while (p6(a + 1) <= n) ++a;
  while (a && p6(a) > n) --a;
  return (u32)a;
}

void find3_window(u128 T, u32 lim, u32& lo, u32& hi) {
  hi = std::min(iroot6(T), lim);
  auto p6 = [](u64 v) { const u128 v2 = (u128)v * v; return v2 * v2 * v2; };
  const double x = (double)(u64)(T >> 64) * 18446744073709551616.0 + (double)(u64)T;
  u64 l = (u64)std::cbrt(std::sqrt(x / 3.0));
  if (l < 1) l = 1;
  while (l > 1 && p6(l - 1) * 3 >= T) --l;
  while (l <= hi && p6(l) * 3 < T) ++l;
  lo = (u32)l;
}

void expand_cls234(const Cls234Unit& U, std::vector<std::pair<u32, u128>>& jobs, u64& failT) {
  for (u32 i = 0; i < U.nres; ++i)
    for (u64 d = U.res[i]; d <= U.fmax; d += U.mod) {
      if (!d) continue;
      u128 T;
      if (U.regular) {
        const u128 d2 = (u128)d * d, d6 = d2 * d2 * d2;
        const u128 dq = ((d6 - U.dr) >> U.shift) * U.inv;
        if (U.Q <= dq) { ++failT; continue; }
        T = U.Q - dq;
      } else if (!exact_T_cls234(U.B, U.u, U.f, d, T)) {
        ++failT;
        continue;
      }
      if (T == 0) continue;
      jobs.emplace_back((u32)d, T);
    }
}

bool verify_cls234(u64 B, u64 u, u64 f, u64 d, u64 a, u64 b, u64 c) {
  U256 rhs = U256::pow6(u);
  add_in_place(rhs, U256::pow6(f * d));
  U256 core = U256::pow6(a);
  add_in_place(core, U256::pow6(b));
  add_in_place(core, U256::pow6(c));
  mul_small(core, kM42);
  add_in_place(rhs, core);
  return cmp(rhs, U256::pow6(B)) == 0;
}

// ---------------------------------------------------------------------------
// pair recovery / verify
// ---------------------------------------------------------------------------
void exact_pairs(u128 T, u32 hi, std::vector<std::pair<u32, u32>>& out) {
  out.clear();
  auto p6 = [](u64 x) { const u128 x2 = (u128)x * x; return x2 * x2 * x2; };
  u64 b = std::min<u64>(hi, iroot6(T));
  for (u64 a = 1; a <= b; ++a) {
    const u128 pa = p6(a);
    if (pa > T) break;
    while (b >= a && pa + p6(b) > T) --b;
    if (b < a) break;
    if (pa + p6(b) == T) out.emplace_back((u32)a, (u32)b);
  }
}

// the exact pair target of a class-5 hit: T itself, pairs up to min(lim, N)
static PairTarget cls5_target(int N, u64 B, u64 u, u32 lim, u32 d, u32 e, u64 fp) {
  PairTarget t{};
  u128 T;
  if (!exact_T_cls5(B, u, d, e, T) || (u64)T != fp) return t;
  t.T = T;
  t.hi = std::min<u32>(lim, (u32)std::max(N, 0));
  t.ok = true;
  return t;
}

bool solution_record(u64 B, u64 u, u32 cls, u32 d, u32 e, u32 a, u32 b, u32 c, const i64 terms[5], Solution& out) {
  i64 t[5];
  std::copy(terms, terms + 5, t);
  std::sort(t, t + 5);
  for (int i = 0; i < 4; ++i)
    if (t[i] == t[i + 1]) return false;
  if (t[4] >= (i64)B || t[0] < 1) return false;
  i64 g = t[0];
  for (int i = 1; i < 5; ++i) g = gcd_i64(g, t[i]);
  g = gcd_i64(g, (i64)B);
  out = Solution{B, u, cls, d, e, a, b, c, {t[0], t[1], t[2], t[3], t[4]}, g};
  return true;
}

static int cls5_finish(u64 B, u64 u, u32 d, u32 e, u32 a, u32 b, std::vector<Solution>& out) {
  if (!verify_cls5(B, u, d, e, a, b)) return 0;
  const i64 t[5] = {42LL * a, 42LL * b, 21LL * (i64)d, 14LL * (i64)e, (i64)u};
  Solution s;
  if (!solution_record(B, u, 5, d, e, a, b, 0, t, s)) return 0;
  out.push_back(s);
  return 1;
}

int verify_cls5_hit(int N, u64 B, u64 u, u32 lim, u32 d, u32 e, u64 fp, std::vector<Solution>& out) {
  const PairTarget t = cls5_target(N, B, u, lim, d, e, fp);
  if (!t.ok) return 0;
  std::vector<std::pair<u32, u32>> pairs;
  exact_pairs(t.T, t.hi, pairs);
  int found = 0;
  for (auto [a, b] : pairs) found += cls5_finish(B, u, d, e, a, b, out);
  return found;
}

}  // namespace fc5

namespace fc5 {
// the exact pair target of a class 2-4 hit at (d, c3): T - c3^6, pairs b <= min(c3, N)
// (fourcore_find_v4.cu:548)
static PairTarget find3_target(int N, int cls, u64 B, u64 u, u32 d, u32 c, u64 fp) {
  PairTarget t{};
  const u64 f = (u64)free_factor(cls);
  u128 T;
  if (c < 1 || !exact_T_cls234(B, u, f, d, T)) return t;
  const u64 c3 = (u64)c * c * c;
  const u128 c6 = (u128)c3 * c3;
  if (c6 > T || (u64)(T - c6) != fp) return t;
  t.T = T - c6;
  t.hi = std::min<u32>(c, (u32)std::max(N, 0));
  t.ok = true;
  return t;
}

static int find3_finish(int cls, u64 B, u64 u, u32 d, u32 c, u32 a, u32 b, std::vector<Solution>& out) {
  const u64 f = (u64)free_factor(cls);
  if (!verify_cls234(B, u, f, d, a, b, c)) return 0;
  const i64 t[5] = {42LL * a, 42LL * b, 42LL * c, (i64)f * (i64)d, (i64)u};
  Solution s;
  if (!solution_record(B, u, (u32)cls, d, 0, a, b, c, t, s)) return 0;
  out.push_back(s);
  return 1;
}

int verify_find3_hit(int N, int cls, u64 B, u64 u, u32 d, u32 c, u64 fp, std::vector<Solution>& out) {
  const PairTarget t = find3_target(N, cls, B, u, d, c, fp);
  if (!t.ok) return 0;
  std::vector<std::pair<u32, u32>> pairs;
  exact_pairs(t.T, t.hi, pairs);
  int found = 0;
  for (auto [a, b] : pairs) found += find3_finish(cls, B, u, d, c, a, b, out);
  return found;
}

PairTarget verify_target(int N, const VerifyJob& j) {
  if (j.cls == 5) return cls5_target(N, j.B, j.u, j.lim, j.d, j.e, j.fp);
  if (j.cls >= 2 && j.cls <= 4) return find3_target(N, j.cls, j.B, j.u, j.d, j.e, j.fp);
  return PairTarget{};
}

int verify_finish(const VerifyJob& j, u32 a, u32 b, std::vector<Solution>& out) {
  if (j.cls == 5) return cls5_finish(j.B, j.u, j.d, j.e, a, b, out);
  if (j.cls >= 2 && j.cls <= 4) return find3_finish(j.cls, j.B, j.u, j.d, j.e, a, b, out);
  return 0;
}

void verify_jobs(int N, const std::vector<VerifyJob>& jobs, bool stop_first, std::vector<Solution>& out) {
  std::vector<std::vector<Solution>> per(jobs.size());
#pragma omp parallel for schedule(dynamic, 1) if (jobs.size() > 1)
  for (long i = 0; i < (long)jobs.size(); ++i) {
    const VerifyJob& j = jobs[(size_t)i];
    if (j.cls == 5)
      verify_cls5_hit(N, j.B, j.u, j.lim, j.d, j.e, j.fp, per[(size_t)i]);
    else if (j.cls >= 2 && j.cls <= 4)
      verify_find3_hit(N, j.cls, j.B, j.u, j.d, j.e, j.fp, per[(size_t)i]);
  }
  for (auto& v : per) {
    out.insert(out.end(), v.begin(), v.end());
    if (stop_first && !v.empty()) break;
  }
}
}  // namespace fc5

// ---------------------------------------------------------------------------
// (6,1,7) class-5 twin
// ---------------------------------------------------------------------------
namespace fc5 {

static u128 pow6_128(u64 x) {
  const u128 x2 = (u128)x * x;
  return x2 * x2 * x2;
}

static std::vector<i64> lift(const std::vector<i64>& res, i64 m, i64 maxv) {
  std::vector<i64> v;
  for (i64 r : res)
    for (i64 x = (r % m) ? r % m : m; x <= maxv; x += m) v.push_back(x);
  std::sort(v.begin(), v.end());
  return v;
}

void twin_cands(u64 B, double u_lo, double u_hi, std::vector<TwinCandRaw>& out) {
  const i64 Bi = (i64)B;
  const i64 lo = std::max<i64>(1, (i64)(u_lo * (double)B));
  const i64 hi = std::min<i64>(Bi - 1, (i64)(u_hi * (double)B));
  // the three role terms: (21w')^6 = B^6 (2^6), (14u')^6 = B^6 (3^6), (6v')^6 = B^6 (7^6)
  const auto wres = sixth_roots(pow6_mod(Bi, kM2) * inv_mod(pow6_mod(21, kM2), kM2) % kM2, kM2);
  const auto ures = sixth_roots(pow6_mod(Bi, kM3) * inv_mod(pow6_mod(14, kM3), kM3) % kM3, kM3);
  const auto vres = sixth_roots(pow6_mod(Bi, kM7) * inv_mod(pow6_mod(6, kM7), kM7) % kM7, kM7);
  const auto uv = lift(ures, kM3, (Bi - 1) / 14);
  const auto vv = lift(vres, kM7, (Bi - 1) / 6);
  const U256 B6 = U256::pow6(B);
  for (i64 up : uv) {
    const i64 u = 14 * up;
    if (u < lo || u > hi) continue;
    U256 bu = B6;
    if (sub_in_place(bu, U256::pow6((u64)u))) continue;
    for (i64 vp : vv) {
      U256 base = bu;
      if (sub_in_place(base, U256::pow6((u64)(6 * vp))) || base.is_zero()) continue;
      if (!base.fits128()) continue;  // B^6 >= 2^128: outside the reference's i128 range
      TwinCandRaw C{};
      C.B = B;
      C.u = (u64)u;
      C.w = (u64)(6 * vp);
      C.base = base.lo128();
      C.lim = (u32)((B - 1) / 42);
      C.fmax1 = (u32)((B - 1) / 21);
      C.mod1 = (u32)kM2;
      C.nres1 = (u32)std::min<size_t>(wres.size(), 8);
      for (u32 i = 0; i < C.nres1; ++i) C.res[i] = (u32)wres[i];
      out.push_back(C);
    }
  }
}

bool pack_twin(const TwinCandRaw& in, TwinCand& C, std::string& why) {
  C = TwinCand{};
  if (in.mod1 != 64) { why = "grid modulus must be 64"; return false; }
  if (in.nres1 < 1 || in.nres1 > 8) { why = "1..8 grid residues required"; return false; }
  if (in.base >> 127) { why = "base must be below 2^127"; return false; }
  if (in.lim >= 65536) { why = "lim must be below 65536"; return false; }
  if (in.fmax1 >= (1u << 21)) { why = "fmax1 must be below 2^21"; return false; }
  C.B = in.B;
  C.u = in.u;
  C.w = in.w;
  C.base = in.base;
  C.lim = in.lim;
  C.fmax1 = in.fmax1;
  C.nres = in.nres1;
  for (u32 i = 0; i < in.nres1; ++i) {
    if (in.res[i] >= 64) { why = "grid residue must be below 64"; return false; }
    C.res[i] = in.res[i];
    const u32 t = (u32)pow6_mod((i64)in.res[i], 64);
    if (i == 0) C.td = t;
    else if (t != C.td) { why = "grid residues with different sixth powers mod 64"; return false; }
  }
  const u64 m21 = 85766121ULL;  // 21^6
  if ((u64)(in.base % (u128)kM42) != m21 * C.td) {
    why = "base - (21d)^6 is not a multiple of 42^6";
    return false;
  }
  C.Q = in.base / (u128)kM42;
  return true;
}

PairTarget twin_target(int N, const TwinCand& C, u32 d, u32 c4, u32 c3, u64 fp) {
  PairTarget t{};
  const u128 g6 = pow6_128(21ULL * d);
  if (g6 > C.base) return t;
  const u128 rhs = C.base - g6;
  if (rhs % (u128)kM42) return t;
  const u128 T = rhs / (u128)kM42, a6 = pow6_128(c4), b6 = pow6_128(c3);
  if (a6 > T || b6 > T - a6) return t;
  const u128 T2 = T - a6 - b6;
  if ((u64)T2 != fp) return t;
  t.T = T2;
  t.hi = (u32)std::max(N, 0);
  t.ok = true;
  return t;
}

int twin_finish(const TwinCand& C, u32 d, u32 c4, u32 c3, u32 a, u32 b, std::vector<TwinSolution>& out) {
  i64 t[7] = {42LL * a, 42LL * b, 42LL * c3, 42LL * c4, 21LL * (i64)d, (i64)C.u, (i64)C.w};
  std::sort(t, t + 7);
  bool bad = t[0] < 1 || t[6] >= (i64)C.B;
  for (int i = 0; i < 6; ++i) bad |= t[i] == t[i + 1];
  if (bad) return 0;
  U256 lhs = U256::of(0);
  for (int i = 0; i < 7; ++i) add_in_place(lhs, U256::pow6((u64)t[i]));
  if (cmp(lhs, U256::pow6(C.B)) != 0) return 0;
  i64 g = t[0];
  for (int i = 1; i < 7; ++i) g = std::gcd(g, t[i]);
  g = std::gcd(g, (i64)C.B);
  TwinSolution S{C.B, C.u, C.w, d, c4, c3, a, b, {}, g};
  for (int i = 0; i < 7; ++i) S.terms[i] = t[i];
  out.push_back(S);
  return 1;
}

int verify_twin_hit(int N, const TwinCand& C, u32 d, u32 c4, u32 c3, u64 fp, std::vector<TwinSolution>& out) {
  const PairTarget t = twin_target(N, C, d, c4, c3, fp);
  if (!t.ok) return 0;
  std::vector<std::pair<u32, u32>> pairs;
  exact_pairs(t.T, t.hi, pairs);
  int found = 0;
  for (auto [a, b] : pairs) found += twin_finish(C, d, c4, c3, a, b, out);
  return found;
}

void verify_twin_hits(int N, const std::vector<TwinCand>& cands, const std::vector<TwinHitIn>& hits,
                      std::vector<std::vector<TwinSolution>>& per) {
  per.assign(hits.size(), {});
#pragma omp parallel for schedule(dynamic, 1) if (hits.size() > 1)
  for (long i = 0; i < (long)hits.size(); ++i) {
    const TwinHitIn& h = hits[(size_t)i];
    if (h.cand < cands.size()) verify_twin_hit(N, cands[h.cand], h.d, h.c4, h.c3, h.fp, per[(size_t)i]);
  }
}

void verify_targets(int N, const std::vector<VerifyJob>& jobs, std::vector<PairTarget>& out) {
  out.assign(jobs.size(), PairTarget{});
#pragma omp parallel for schedule(dynamic, 64) if (jobs.size() > 256)
  for (long i = 0; i < (long)jobs.size(); ++i) out[(size_t)i] = verify_target(N, jobs[(size_t)i]);
}

void verify_finish_all(const std::vector<VerifyJob>& jobs, const std::vector<std::vector<std::pair<u32, u32>>>& pairs,
                       bool stop_first, std::vector<Solution>& out) {
  for (size_t i = 0; i < jobs.size() && i < pairs.size(); ++i) {
    int found = 0;
    for (auto [a, b] : pairs[i]) found += verify_finish(jobs[i], a, b, out);
    if (stop_first && found) break;
  }
}

void twin_targets(int N, const std::vector<TwinCand>& cands, const std::vector<TwinHitIn>& hits,
                  std::vector<PairTarget>& out) {
  out.assign(hits.size(), PairTarget{});
#pragma omp parallel for schedule(dynamic, 64) if (hits.size() > 256)
  for (long i = 0; i < (long)hits.size(); ++i) {
    const TwinHitIn& h = hits[(size_t)i];
    if (h.cand < cands.size()) out[(size_t)i] = twin_target(N, cands[h.cand], h.d, h.c4, h.c3, h.fp);
  }
}

void pack_units(const u32* cls, const u64* bu, size_t n, std::vector<Cls5Unit>& u5, std::vector<char>& ok5,
                std::vector<Cls234Unit>& u3, std::vector<char>& ok3, std::vector<int>& slot) {
  slot.assign(n, -1);
  size_t n5 = 0, n3 = 0;
  for (size_t i = 0; i < n; ++i) {
    const u32 c = cls ? cls[i] : 5u;
    if (c == 5) slot[i] = (int)n5++;
    else if (c >= 2 && c <= 4) slot[i] = (int)n3++;
  }
  u5.assign(n5, Cls5Unit{});
  ok5.assign(n5, 0);
  u3.assign(n3, Cls234Unit{});
  ok3.assign(n3, 0);
#pragma omp parallel for schedule(dynamic, 256) if (n > 4096)
  for (long i = 0; i < (long)n; ++i) {
    const int s = slot[(size_t)i];
    if (s < 0) continue;
    const u32 c = cls ? cls[i] : 5u;
    const u64 B = bu[2 * i], u = bu[2 * i + 1];
    if (c == 5) ok5[(size_t)s] = pack_cls5(B, u, u5[(size_t)s]) ? 1 : 0;
    else ok3[(size_t)s] = pack_cls234((int)c, B, u, u3[(size_t)s]) ? 1 : 0;
  }
}

}  // namespace fc5
