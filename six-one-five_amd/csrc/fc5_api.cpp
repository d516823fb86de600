// fc5_api.cpp -- the C ABI of include/fc5.h over the host core and kernels.
#include "fc5.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fc5_build.hpp"
#include "fc5_host.hpp"
#include "fc5_kernels.hpp"

using namespace fc5;

namespace {

constexpr u64 kGateModU = 124488;
constexpr u32 kGateWordsU = (124488 + 31) / 32;

struct DeviceTable {
  StoreHeader hdr{};
  u64* cells = nullptr;
  XorDev dev{};
  bool present = false;
};

void free_table(DeviceTable& t) {
  if (t.cells) (void)hipFree(t.cells);
  t = DeviceTable{};
}

u64 inv_mod_2_64(u64 a) {  // a odd; Newton iteration doubles correct bits
  u64 x = a;
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}

}  // namespace

struct fc5_band {
  std::vector<Cls5Unit> units;     // packed units (skipped ones removed)
  std::vector<u32> uid;            // caller index of each packed unit
  std::vector<u32> irregular;      // indices into units: per-point 192-bit path
  DevUnit5* d_units = nullptr;
  u32 n_dev = 0, n_items = 0;
  PlainUnit* d_plain = nullptr;    // every packed unit, reference shape
  u32 n_plain = 0;
  u32 max_fmax2 = 0, max_lim = 0;
  size_t n_in = 0, n_skipped = 0;
};

struct fc5_ctx {
  int device = 0;
  int num_cu = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // options
  u64 hit_cap = 1u << 20;
  int use_gate = 1, verify = 1, stop_first = 0, blocks_per_cu = 0;
  u32 load_policy = 1;  // cache policy of the xor cell gathers (nontemporal: no MALL allocation)
  // tables
  DeviceTable main, side;
  XorTable host_main;  // host copy when small enough (save, host verify)
  bool host_main_valid = false;
  // gate
  u32* d_gate = nullptr;
  u32* d_gate_ones = nullptr;
  GateWords gw{};
  // d table
  u32* d_ndg = nullptr;
  u32 ndg_K = 0, ndg_dmax = 0;
  // run buffers
  HitRec* d_hits = nullptr;
  u64 d_hits_cap = 0;
  Counters* d_ctr = nullptr;  // [0] batched, [1] plain
  u64* d_nhits = nullptr;
  u32* d_item_ctr = nullptr;
  // last results
  std::vector<fc5_hit> hits;
  std::vector<fc5_solution> sols;
  PairRecover pr;
  int pr_N = -1;
};

#define FAIL(ctx, code, msg)  \
  do {                        \
    (ctx)->err = (msg);       \
    return (code);            \
  } while (0)

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      (ctx)->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr;   \
      return FC5_E_HIP;                                                                \
    }                                                                                  \
  } while (0)

static void build_gate_host(GateWords& gw, std::vector<u32>& bitmap) {
  // pair-sum sets of sixth powers mod 504 and 247, 0 included
  // (fourcore_cls5_gpu_v4.cu:124-137)
  std::memset(&gw, 0, sizeof gw);
  auto p6 = [](u32 x, u32 m) { u64 a = x % m, a2 = a * a % m; return (u32)(a2 * a2 % m * a2 % m); };
  for (u32 i = 0; i < 504; ++i)
    for (u32 j = i; j < 504; ++j) {
      const u32 s = (p6(i, 504) + p6(j, 504)) % 504;
      gw.g504[s >> 6] |= 1ULL << (s & 63);
    }
  for (u32 i = 0; i < 247; ++i)
    for (u32 j = i; j < 247; ++j) {
      const u32 s = (p6(i, 247) + p6(j, 247)) % 247;
      gw.g247[s >> 6] |= 1ULL << (s & 63);
    }
  bitmap.assign(kGateWordsU, 0);
  for (u32 t = 0; t < kGateModU; ++t) {
    const u32 a = t % 504, b = t % 247;
    if (((gw.g504[a >> 6] >> (a & 63)) & 1) && ((gw.g247[b >> 6] >> (b & 63)) & 1))
      bitmap[t >> 5] |= 1u << (t & 31);
  }
}

extern "C" int fc5_abi_version(void) { return FC5_ABI_VERSION; }

extern "C" int fc5_ctx_create(int device, fc5_ctx** out) {
  if (!out) return FC5_E_ARG;
  *out = nullptr;
  auto ctx = std::make_unique<fc5_ctx>();
  ctx->device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return FC5_E_HIP;
  if (hipSetDevice(device) != hipSuccess) return FC5_E_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FC5_E_HIP;
  ctx->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return FC5_E_HIP;
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
    return FC5_E_HIP;
  std::vector<u32> bitmap;
  build_gate_host(ctx->gw, bitmap);
  std::vector<u32> ones(kGateWordsU, ~0u);
  if (hipMalloc(&ctx->d_gate, kGateWordsU * 4) != hipSuccess ||
      hipMalloc(&ctx->d_gate_ones, kGateWordsU * 4) != hipSuccess ||
      hipMalloc(&ctx->d_ctr, 4 * sizeof(Counters)) != hipSuccess ||
      hipMalloc(&ctx->d_nhits, 8) != hipSuccess || hipMalloc(&ctx->d_item_ctr, 8) != hipSuccess)
    return FC5_E_NOMEM;
  if (hipMemcpy(ctx->d_gate, bitmap.data(), kGateWordsU * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ctx->d_gate_ones, ones.data(), kGateWordsU * 4, hipMemcpyHostToDevice) != hipSuccess)
    return FC5_E_HIP;
  *out = ctx.release();
  return FC5_OK;
}

extern "C" void fc5_ctx_destroy(fc5_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_table(ctx->main);
  free_table(ctx->side);
  for (void* p : {(void*)ctx->d_gate, (void*)ctx->d_gate_ones, (void*)ctx->d_ndg, (void*)ctx->d_hits,
                  (void*)ctx->d_ctr, (void*)ctx->d_nhits, (void*)ctx->d_item_ctr})
    if (p) (void)hipFree(p);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* fc5_last_error(const fc5_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int fc5_set_option(fc5_ctx* ctx, const char* key, int64_t v) {
  if (!ctx || !key) return FC5_E_ARG;
  const std::string k = key;
  if (k == "hit_cap") {
    if (v < 1 || v > (int64_t)1 << 32) FAIL(ctx, FC5_E_ARG, "hit_cap out of range");
    ctx->hit_cap = (u64)v;
  } else if (k == "use_gate") {
    ctx->use_gate = v ? 1 : 0;
  } else if (k == "verify") {
    ctx->verify = v ? 1 : 0;
  } else if (k == "stop_first") {
    ctx->stop_first = v ? 1 : 0;
  } else if (k == "load_policy") {
    if (v < 0 || v > 3) FAIL(ctx, FC5_E_ARG, "load_policy must be 0..3");
    ctx->load_policy = (u32)v;
    ctx->main.dev.policy = ctx->load_policy;
    ctx->side.dev.policy = ctx->load_policy;
  } else if (k == "blocks_per_cu") {
    if (v < 0 || v > 64) FAIL(ctx, FC5_E_ARG, "blocks_per_cu out of range");
    ctx->blocks_per_cu = (int)v;
  } else {
    FAIL(ctx, FC5_E_ARG, "unknown option " + k);
  }
  return FC5_OK;
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
static int upload_table(fc5_ctx* ctx, const StoreHeader& h, const u8* packed, size_t len,
                        DeviceTable& dst) {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  free_table(dst);
  dst.hdr = h;
  const size_t want = packed_bytes(h.m_cells, h.r);
  if (len < want) FAIL(ctx, FC5_E_FORMAT, "packed length shorter than the header implies");
  if (h.n_keys == 0 || h.m_cells < 3) {  // empty filter: every query answers "no"
    dst.present = true;
    dst.dev = XorDev{};
    return FC5_OK;
  }
  const size_t words = (want + 3) / 4 + 4;  // +16 bytes of zero pad for 3-word reads
  u32* staging = nullptr;
  if (hipMalloc(&staging, words * 4) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "device staging allocation");
  if (hipMalloc(&dst.cells, h.m_cells * 8) != hipSuccess) {
    (void)hipFree(staging);
    FAIL(ctx, FC5_E_NOMEM, "device cell allocation");
  }
  int rc = FC5_OK;
  if (hipMemsetAsync(staging, 0, words * 4, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(staging, packed, want, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      launch_expand_cells(staging, h.m_cells, h.r, dst.cells, ctx->stream) != 0 ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ctx->err = "table upload failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(staging);
  if (rc != FC5_OK) { free_table(dst); return rc; }
  dst.dev.cells = dst.cells;
  dst.dev.block = h.m_cells / 3;
  dst.dev.mask = fp_mask(h.r);
  dst.dev.seed = h.mix_seed;
  dst.dev.md = make_moddiv(dst.dev.block);
  dst.dev.present = dst.dev.block ? 1u : 0u;
  dst.dev.policy = ctx->load_policy;
  dst.present = true;
  return FC5_OK;
}

static constexpr size_t kKeepHostTable = (size_t)4 << 30;

static int adopt_main(fc5_ctx* ctx, XorTable&& t) {
  const int rc = upload_table(ctx, t.hdr, t.packed.data(), t.packed.size(), ctx->main);
  if (rc != FC5_OK) return rc;
  ctx->host_main_valid = t.packed.size() <= kKeepHostTable;
  if (ctx->host_main_valid) ctx->host_main = std::move(t);
  else ctx->host_main = XorTable{};
  ctx->pr_N = -1;
  return FC5_OK;
}

extern "C" int fc5_table_load(fc5_ctx* ctx, const char* path, int expected_N) {
  if (!ctx || !path) return FC5_E_ARG;
  XorTable t;
  std::string msg;
  const int st = load_bin(path, expected_N, t, msg);
  switch (st) {
    case kLoadOk: break;
    case kLoadOpen:
    case kLoadShort: FAIL(ctx, FC5_E_IO, msg);
    case kLoadNMismatch: FAIL(ctx, FC5_E_NMISMATCH, msg);
    default: FAIL(ctx, FC5_E_FORMAT, msg);
  }
  const int fpr = fpr_smoke(t);
  if (fpr < 0 || fpr > 16)
    FAIL(ctx, FC5_E_FPR, "xor FPR smoke too high (" + std::to_string(fpr) + ") - refuse to search");
  return adopt_main(ctx, std::move(t));
}

extern "C" int fc5_table_upload(fc5_ctx* ctx, const fc5_store_header* hdr, const uint8_t* packed,
                                size_t len) {
  if (!ctx || !hdr || (!packed && len)) return FC5_E_ARG;
  XorTable t;
  std::memcpy(&t.hdr, hdr, sizeof t.hdr);
  if (t.hdr.magic != 0x53353136u || t.hdr.version != 2 || t.hdr.kind != 1 || t.hdr.r == 0 || t.hdr.r > 64)
    FAIL(ctx, FC5_E_FORMAT, "bad header");
  t.packed.assign(packed, packed + len);
  return adopt_main(ctx, std::move(t));
}

extern "C" int fc5_table_build(fc5_ctx* ctx, int N, int r, const uint64_t* extra, size_t n_extra,
                               int on_gpu) {
  if (!ctx || N < 1 || r < 1 || r > 64 || (!extra && n_extra)) return FC5_E_ARG;
  if (on_gpu) {
    HIPCHK(ctx, hipSetDevice(ctx->device));
    free_table(ctx->main);
    ctx->host_main = XorTable{};
    ctx->host_main_valid = false;
    GpuBuilt g;
    std::string err;
    if (gpu_build_pairs(N, (u32)r, extra, n_extra, ctx->stream, g, err) != 0)
      FAIL(ctx, err.find("stalled") != std::string::npos ? FC5_E_BUILD : FC5_E_HIP, err);
    DeviceTable& t = ctx->main;
    std::memcpy(&t.hdr, &g.hdr, sizeof t.hdr);
    t.cells = g.cells;
    t.dev.cells = g.cells;
    t.dev.block = g.hdr.m_cells / 3;
    t.dev.mask = fp_mask((u32)r);
    t.dev.seed = g.hdr.mix_seed;
    t.dev.md = make_moddiv(t.dev.block);
    t.dev.present = t.dev.block ? 1u : 0u;
    t.dev.policy = ctx->load_policy;
    t.present = true;
    ctx->pr_N = -1;
    return FC5_OK;
  }
  std::vector<u64> ex(extra, extra + n_extra);
  XorTable t;
  std::string msg;
  if (!build_pairs_cpu(N, (u32)r, ex, t, msg)) FAIL(ctx, FC5_E_BUILD, msg);
  return adopt_main(ctx, std::move(t));
}

extern "C" int fc5_table_info(const fc5_ctx* ctx, fc5_store_header* out) {
  if (!ctx || !out) return FC5_E_ARG;
  if (!ctx->main.present) return FC5_E_STATE;
  std::memcpy(out, &ctx->main.hdr, sizeof *out);
  return FC5_OK;
}

// pack the table to the file layout (host copy, or the device cells re-packed)
static int table_packed(fc5_ctx* ctx, XorTable& t, bool want_bytes) {
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  t.hdr = ctx->main.hdr;
  if (!want_bytes) return FC5_OK;
  if (ctx->host_main_valid) {
    t.packed = ctx->host_main.packed;
    return FC5_OK;
  }
  const u64 m = t.hdr.m_cells;
  const size_t want = packed_bytes(m, t.hdr.r);
  t.packed.assign(want, 0);
  if (!ctx->main.cells || !want) return FC5_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t room = std::max<size_t>(want, (size_t)((m + 7) / 8) * t.hdr.r);
  u8* d = nullptr;
  if (hipMalloc(&d, room) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "pack buffer");
  int rc = FC5_OK;
  if (hipMemsetAsync(d, 0, room, ctx->stream) != hipSuccess ||
      gpu_pack_cells(ctx->main.cells, m, t.hdr.r, d, ctx->stream) != 0 ||
      hipMemcpyAsync(t.packed.data(), d, want, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ctx->err = "table pack failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(d);
  return rc;
}

extern "C" int fc5_table_save(fc5_ctx* ctx, const char* path) {
  if (!ctx || !path) return FC5_E_ARG;
  XorTable t;
  const int rc = table_packed(ctx, t, true);
  if (rc != FC5_OK) return rc;
  std::string msg;
  if (!save_bin(path, t, msg)) FAIL(ctx, FC5_E_IO, msg);
  return FC5_OK;
}

extern "C" int fc5_table_download(fc5_ctx* ctx, fc5_store_header* hdr, uint8_t* packed, size_t cap, size_t* len) {
  if (!ctx || !len) return FC5_E_ARG;
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  const size_t want = packed_bytes(ctx->main.hdr.m_cells, ctx->main.hdr.r);
  *len = want;
  if (hdr) std::memcpy(hdr, &ctx->main.hdr, sizeof *hdr);
  if (!packed) return FC5_OK;
  if (cap < want) FAIL(ctx, FC5_E_ARG, "buffer too small");
  XorTable t;
  const int rc = table_packed(ctx, t, true);
  if (rc != FC5_OK) return rc;
  std::memcpy(packed, t.packed.data(), want);
  return FC5_OK;
}

static int device_query(fc5_ctx* ctx, const XorDev& x, const uint64_t* keys, size_t n, uint8_t* out) {
  if (n == 0) return FC5_OK;
  u64* dk = nullptr;
  u8* dout = nullptr;
  if (hipMalloc(&dk, n * 8) != hipSuccess || hipMalloc(&dout, n) != hipSuccess) {
    if (dk) (void)hipFree(dk);
    FAIL(ctx, FC5_E_NOMEM, "query buffers");
  }
  int rc = FC5_OK;
  if (hipMemcpyAsync(dk, keys, n * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      launch_query(x, dk, n, dout, ctx->stream) != 0 ||
      hipMemcpyAsync(out, dout, n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    ctx->err = "device query failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(dk);
  (void)hipFree(dout);
  return rc;
}

extern "C" int fc5_table_query(fc5_ctx* ctx, const uint64_t* keys, size_t n, uint8_t* out) {
  if (!ctx || (!keys && n) || (!out && n)) return FC5_E_ARG;
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return device_query(ctx, ctx->main.dev, keys, n, out);
}

extern "C" int fc5_table_fpr_smoke(fc5_ctx* ctx, int trials, int* hits) {
  if (!ctx || !hits || trials < 0) return FC5_E_ARG;
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  std::vector<u64> keys((size_t)trials);
  u64 x = 0x123456789abcdef0ULL;
  for (auto& k : keys) { x = x * kPhi + 1; k = x; }
  std::vector<u8> out((size_t)trials);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int rc = device_query(ctx, ctx->main.dev, keys.data(), keys.size(), out.data());
  if (rc != FC5_OK) return rc;
  int h = 0;
  for (u8 v : out) h += v;
  *hits = h;
  return FC5_OK;
}

extern "C" int fc5_side_build(fc5_ctx* ctx, const uint64_t* keys, size_t n, int r) {
  if (!ctx || (!keys && n) || r < 1 || r > 64) return FC5_E_ARG;
  std::vector<u64> k(keys, keys + n);
  XorTable t;
  if (!build_xor_cpu(k, (u32)r, 0x51DE51DE51DEULL, 64, t)) FAIL(ctx, FC5_E_BUILD, "side table build stalled");
  return upload_table(ctx, t.hdr, t.packed.data(), t.packed.size(), ctx->side);
}

extern "C" int fc5_side_clear(fc5_ctx* ctx) {
  if (!ctx) return FC5_E_ARG;
  free_table(ctx->side);
  return FC5_OK;
}

// ---------------------------------------------------------------------------
// bands
// ---------------------------------------------------------------------------
static int ensure_ndg(fc5_ctx* ctx, u32 dmax) {
  if (ctx->d_ndg && ctx->ndg_dmax >= dmax) return FC5_OK;
  const u32 K = dmax / 64 + 1;
  std::vector<u32> h((size_t)64 * K);
  for (u32 r = 0; r < 64; ++r)
    for (u32 k = 0; k < K; ++k) {
      const u64 d = (u64)r + 64ull * k;
      const u64 d3 = d * d * d;
      const u128 dq = ((u128)d3 * d3) >> 6;
      h[(size_t)r * K + k] = (u32)(kGateModU - (u64)(dq % kGateModU));
    }
  if (ctx->d_ndg) (void)hipFree(ctx->d_ndg);
  ctx->d_ndg = nullptr;
  if (hipMalloc(&ctx->d_ndg, h.size() * 4) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "d table");
  HIPCHK(ctx, hipMemcpy(ctx->d_ndg, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  ctx->ndg_K = K;
  ctx->ndg_dmax = 64 * K - 1;
  return FC5_OK;
}

static int ensure_hits(fc5_ctx* ctx) {
  if (ctx->d_hits && ctx->d_hits_cap == ctx->hit_cap) return FC5_OK;
  if (ctx->d_hits) (void)hipFree(ctx->d_hits);
  ctx->d_hits = nullptr;
  if (hipMalloc(&ctx->d_hits, ctx->hit_cap * sizeof(HitRec)) != hipSuccess)
    FAIL(ctx, FC5_E_NOMEM, "hit buffer");
  ctx->d_hits_cap = ctx->hit_cap;
  return FC5_OK;
}

extern "C" int fc5_band_create(fc5_ctx* ctx, const fc5_unit* units, size_t n, fc5_band** out) {
  if (!ctx || !out || (!units && n)) return FC5_E_ARG;
  if (n > 0xffffffffull) FAIL(ctx, FC5_E_ARG, "too many units in one band");
  *out = nullptr;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto b = std::make_unique<fc5_band>();
  b->n_in = n;
  std::vector<DevUnit5> dev;
  std::vector<PlainUnit> plain;
  u32 item = 0;
  for (size_t i = 0; i < n; ++i) {
    Cls5Unit U;
    if (!pack_cls5(units[i].B, units[i].u, U)) { ++b->n_skipped; continue; }
    const u32 idx = (u32)b->units.size();
    b->max_fmax2 = std::max(b->max_fmax2, U.fmax2);
    b->max_lim = std::max(b->max_lim, U.lim);
    PlainUnit P{};
    P.b0 = U.base.l[0];
    P.b1 = U.base.l[1];
    P.b2 = U.base.l[2];
    P.fmax1 = U.fmax1;
    P.fmax2 = U.fmax2;
    P.mod1 = U.mod1;
    P.mod2 = U.mod2;
    P.nres1 = U.n1;
    P.nres2 = U.n2;
    P.uid = (u32)i;
    for (u32 k = 0; k < U.n1; ++k) P.res[k] = U.res1[k];
    for (u32 k = 0; k < U.n2; ++k) P.res[U.n1 + k] = U.res2[k];
    plain.push_back(P);
    if (U.regular) {
      DevUnit5 D{};
      D.Q_lo = (u64)U.Q;
      D.Q_hi = (u64)(U.Q >> 64);
      D.fmax1 = U.fmax1;
      D.fmax2 = U.fmax2;
      D.rows = 6u * (U.fmax1 / 729u + 1u);
      D.item_base = item;
      item += (D.rows + kRowsPerItem - 1) / kRowsPerItem;
      D.te = U.te;
      D.td = U.td;
      D.uid = (u32)i;
      D.band = 0;
      for (int k = 0; k < 6; ++k) D.res1[k] = (u16)U.res1[k];
      for (int k = 0; k < 4; ++k) D.res2[k] = (u8)U.res2[k];
      dev.push_back(D);
    } else {
      b->irregular.push_back(idx);
    }
    b->units.push_back(U);
    b->uid.push_back((u32)i);
  }
  b->n_dev = (u32)dev.size();
  b->n_items = item;
  b->n_plain = (u32)plain.size();
  if (!dev.empty()) {
    if (hipMalloc(&b->d_units, dev.size() * sizeof(DevUnit5)) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "units");
    HIPCHK(ctx, hipMemcpy(b->d_units, dev.data(), dev.size() * sizeof(DevUnit5), hipMemcpyHostToDevice));
  }
  if (!plain.empty()) {
    if (hipMalloc(&b->d_plain, plain.size() * sizeof(PlainUnit)) != hipSuccess) {
      if (b->d_units) (void)hipFree(b->d_units);
      FAIL(ctx, FC5_E_NOMEM, "plain units");
    }
    HIPCHK(ctx, hipMemcpy(b->d_plain, plain.data(), plain.size() * sizeof(PlainUnit), hipMemcpyHostToDevice));
  }
  const int rc = ensure_ndg(ctx, b->max_fmax2);
  if (rc != FC5_OK) {
    fc5_band_destroy(ctx, b.release());
    return rc;
  }
  *out = b.release();
  return FC5_OK;
}

extern "C" int fc5_band_destroy(fc5_ctx* ctx, fc5_band* band) {
  if (!band) return FC5_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  if (band->d_units) (void)hipFree(band->d_units);
  if (band->d_plain) (void)hipFree(band->d_plain);
  delete band;
  return FC5_OK;
}

static PlainArgs plain_args(fc5_ctx* ctx) {
  PlainArgs p{};
  p.use_gate = ctx->use_gate;
  p.gate = ctx->gw;
  p.main = ctx->main.dev;
  p.side = ctx->side.present ? ctx->side.dev : XorDev{};
  p.inv216 = inv_mod_2_64(85766121ULL);
  const u64 m504 = 504ULL * kM42;
  p.c_m504 = (u64)(((u128)1 << 64) % m504);
  p.c_m247 = (u64)(((u128)1 << 64) % 247u);
  // inverse of 42^6 mod 247
  const u64 p6 = (u64)pow6_mod(42, 247);
  u64 inv = 1;
  for (u64 x = 1; x < 247; ++x)
    if (p6 * x % 247 == 1) { inv = x; break; }
  p.inv42_247 = inv;
  p.hits = ctx->d_hits;
  p.hit_cap = ctx->hit_cap;
  p.nhits = ctx->d_nhits;
  p.ctr = ctx->d_ctr + 1;
  return p;
}

// reference chunking: 16 units per synchronous launch (fourcore_cls5_gpu_v4.cu:468-501)
static int run_plain_units(fc5_ctx* ctx, fc5_band* b, const std::vector<u32>& which, bool sync_each) {
  constexpr int kChunk = 16;
  PlainArgs p = plain_args(ctx);
  PlainUnit* d_chunk = nullptr;
  const bool all = which.empty();
  const size_t total = all ? b->n_plain : which.size();
  if (!all) {
    if (hipMalloc(&d_chunk, kChunk * sizeof(PlainUnit)) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "chunk");
  }
  int rc = FC5_OK;
  for (size_t off = 0; off < total && rc == FC5_OK; off += kChunk) {
    const int n = (int)std::min<size_t>(kChunk, total - off);
    u32 ymax = 1;
    for (int i = 0; i < n; ++i) {
      const Cls5Unit& U = b->units[all ? off + i : which[off + i]];
      ymax = std::max(ymax, U.n1 * (U.fmax1 / U.mod1 + 1));
    }
    if (all) {
      p.units = b->d_plain + off;
    } else {
      // gather the selected units (the plain array holds every packed unit in band order)
      for (int i = 0; i < n; ++i)
        if (hipMemcpyAsync(d_chunk + i, b->d_plain + which[off + i], sizeof(PlainUnit),
                           hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
          rc = FC5_E_HIP;
      p.units = d_chunk;
    }
    p.n_units = n;
    if (launch_cls5_plain(p, n, (int)ymax, ctx->stream) != 0) { ctx->err = "plain launch"; rc = FC5_E_HIP; }
    if (sync_each && hipStreamSynchronize(ctx->stream) != hipSuccess) { ctx->err = "plain sync"; rc = FC5_E_HIP; }
  }
  if (d_chunk) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_chunk);
  }
  return rc;
}

extern "C" int fc5_band_run(fc5_ctx* ctx, fc5_band* b, int mode, fc5_result* res) {
  if (!ctx || !b || !res) return FC5_E_ARG;
  if (mode != FC5_MODE_PLAIN && mode != FC5_MODE_BATCHED) FAIL(ctx, FC5_E_ARG, "unknown mode");
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no pair table loaded");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_hits(ctx);
  if (rc != FC5_OK) return rc;
  std::memset(res, 0, sizeof *res);
  ctx->hits.clear();
  ctx->sols.clear();
  HIPCHK(ctx, hipMemsetAsync(ctx->d_ctr, 0, 4 * sizeof(Counters), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_nhits, 0, 8, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_item_ctr, 0, 8, ctx->stream));
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (mode == FC5_MODE_BATCHED) {
    if (b->n_items) {
      BatchedArgs a{};
      a.units = b->d_units;
      a.n_units = b->n_dev;
      a.n_items = b->n_items;
      a.gate = ctx->use_gate ? ctx->d_gate : ctx->d_gate_ones;
      a.ndg = ctx->d_ndg;
      a.ndg_K = ctx->ndg_K;
      a.use_gate = (u32)ctx->use_gate;
      a.main = ctx->main.dev;
      a.side = ctx->side.present ? ctx->side.dev : XorDev{};
      a.hits = ctx->d_hits;
      a.hit_cap = ctx->hit_cap;
      a.nhits = ctx->d_nhits;
      a.ctr = ctx->d_ctr;
      a.item_ctr = ctx->d_item_ctr;
      a.inv729 = inv_mod_2_64(729);
      const int bpc = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : batched_blocks_per_cu();
      const int grid = std::max(1, std::min<int>(ctx->num_cu * bpc, (int)b->n_items));
      if (launch_cls5_batched(a, grid, ctx->stream) != 0) FAIL(ctx, FC5_E_HIP, "batched launch");
    }
    if (!b->irregular.empty()) {
      rc = run_plain_units(ctx, b, b->irregular, false);
      if (rc != FC5_OK) return rc;
    }
  } else {
    rc = run_plain_units(ctx, b, {}, true);
    if (rc != FC5_OK) return rc;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  Counters c[2];
  u64 nh = 0;
  HIPCHK(ctx, hipMemcpyAsync(c, ctx->d_ctr, 2 * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&nh, ctx->d_nhits, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const u64 stored = std::min<u64>(nh, ctx->hit_cap);
  std::vector<HitRec> raw(stored);
  if (stored)
    HIPCHK(ctx, hipMemcpy(raw.data(), ctx->d_hits, stored * sizeof(HitRec), hipMemcpyDeviceToHost));
  res->calls = c[0].calls + c[1].calls;
  res->neg = c[0].neg + c[1].neg;
  res->zero = c[0].zero + c[1].zero;
  res->probes = c[0].probes + c[1].probes;
  res->gated = (c[0].pos - c[0].probes) + c[1].gated;
  res->maybe = c[0].maybe + c[1].maybe;
  res->side_maybe = c[0].side_maybe + c[1].side_maybe;
  res->n_hits = nh;
  res->hits_dropped = nh - stored;
  res->units = b->units.size();
  res->units_skipped = b->n_skipped;
  res->units_irregular = b->irregular.size();
  res->kernel_ms = ms;
  ctx->hits.resize(stored);
  for (u64 i = 0; i < stored; ++i)
    ctx->hits[i] = fc5_hit{raw[i].unit, raw[i].d, raw[i].e, raw[i].flags, raw[i].fp};
  std::sort(ctx->hits.begin(), ctx->hits.end(), [](const fc5_hit& x, const fc5_hit& y) {
    if (x.unit != y.unit) return x.unit < y.unit;
    if (x.e != y.e) return x.e < y.e;
    if (x.d != y.d) return x.d < y.d;
    return x.flags < y.flags;
  });
  if (ctx->verify && !ctx->hits.empty()) {
    const int N = ctx->main.hdr.N ? (int)ctx->main.hdr.N : (int)b->max_lim;
    if (ctx->pr_N != N) {
      ctx->pr.build(N);
      ctx->pr_N = N;
    }
    // caller index -> packed unit
    std::vector<int> pos(b->n_in, -1);
    for (size_t i = 0; i < b->uid.size(); ++i) pos[b->uid[i]] = (int)i;
    std::vector<Solution> sols;
    for (const fc5_hit& h : ctx->hits) {
      if (!(h.flags & 1u)) continue;
      const Cls5Unit& U = b->units[(size_t)pos[h.unit]];
      verify_cls5_hit(ctx->pr, U.B, U.u, U.lim, h.d, h.e, h.fp, sols);
      if (ctx->stop_first && !sols.empty()) break;
    }
    for (const Solution& S : sols) {
      fc5_solution o{};
      o.B = S.B;
      o.u = S.u;
      o.d = S.d;
      o.e = S.e;
      o.a = S.a;
      o.b = S.b;
      for (int k = 0; k < 5; ++k) o.terms[k] = S.terms[k];
      o.gcd = S.gcd;
      ctx->sols.push_back(o);
    }
  }
  res->solutions = ctx->sols.size();
  res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (res->hits_dropped) {
    ctx->err = "hit buffer overflow: " + std::to_string(res->hits_dropped) + " hits dropped";
    return FC5_E_OVERFLOW;
  }
  return FC5_OK;
}

extern "C" int fc5_run_units(fc5_ctx* ctx, const fc5_unit* units, size_t n, int mode, fc5_result* res) {
  fc5_band* b = nullptr;
  int rc = fc5_band_create(ctx, units, n, &b);
  if (rc != FC5_OK) return rc;
  rc = fc5_band_run(ctx, b, mode, res);
  fc5_band_destroy(ctx, b);
  return rc;
}

extern "C" int fc5_last_hits(fc5_ctx* ctx, fc5_hit* out, size_t cap, size_t* n) {
  if (!ctx || !n || (!out && cap)) return FC5_E_ARG;
  *n = ctx->hits.size();
  std::copy_n(ctx->hits.begin(), std::min(cap, ctx->hits.size()), out);
  return FC5_OK;
}

extern "C" int fc5_last_solutions(fc5_ctx* ctx, fc5_solution* out, size_t cap, size_t* n) {
  if (!ctx || !n || (!out && cap)) return FC5_E_ARG;
  *n = ctx->sols.size();
  std::copy_n(ctx->sols.begin(), std::min(cap, ctx->sols.size()), out);
  return FC5_OK;
}

extern "C" int fc5_units_calls(const fc5_unit* units, size_t n, uint64_t* calls) {
  if ((!units && n) || !calls) return FC5_E_ARG;
  u64 c = 0;
  for (size_t i = 0; i < n; ++i) {
    Cls5Unit U;
    if (pack_cls5(units[i].B, units[i].u, U)) c += cls5_calls(U);
  }
  *calls = c;
  return FC5_OK;
}

extern "C" int fc5_emit_units(uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac, double hi_frac,
                              uint32_t* out_cls, uint64_t* out_B, uint64_t* out_u, size_t cap, size_t* n) {
  if (!n || lo == 0 || hi < lo || (cap && (!out_cls || !out_B || !out_u))) return FC5_E_ARG;
  size_t k = 0;
  for (u64 B = lo; B <= hi; ++B) {
    if (!admissible_B((i64)B)) continue;
    for (int cls = 1; cls <= 5; ++cls) {
      if (!(cls_mask & (1u << cls))) continue;
      for (i64 u : unit_candidates((i64)B, cls, lo_frac, hi_frac)) {
        if (k < cap) {
          out_cls[k] = (u32)cls;
          out_B[k] = B;
          out_u[k] = (u64)u;
        }
        ++k;
      }
    }
  }
  *n = k;
  return FC5_OK;
}

extern "C" int fc5_build_pairs_file(int N, int r, const char* path, fc5_store_header* hdr_out) {
  if (N < 1 || r < 1 || r > 64 || !path) return FC5_E_ARG;
  XorTable t;
  std::string msg;
  if (!build_pairs_cpu(N, (u32)r, {}, t, msg)) return FC5_E_BUILD;
  if (!save_bin(path, t, msg)) return FC5_E_IO;
  if (hdr_out) std::memcpy(hdr_out, &t.hdr, sizeof *hdr_out);
  return FC5_OK;
}

// ---------------------------------------------------------------------------
// measurement
// ---------------------------------------------------------------------------
extern "C" int fc5_measure_gathers(fc5_ctx* ctx, uint64_t n, double* probe_ms, double* raw_ms) {
  if (!ctx || !probe_ms || !raw_ms || n == 0) return FC5_E_ARG;
  const u32 policy = ctx->main.dev.policy;
  if (!ctx->main.present || !ctx->main.cells) FAIL(ctx, FC5_E_STATE, "no device table");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned long long* sink = nullptr;
  if (hipMalloc(&sink, 8) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "sink");
  const int grid = ctx->num_cu * 8;
  float a = 0, b = 0;
  int rc = FC5_OK;
  if (hipMemsetAsync(sink, 0, 8, ctx->stream) != hipSuccess ||
      launch_probe_only(ctx->main.dev, n / 16, 1, sink, grid, ctx->stream) != 0 ||  // warm-up
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_probe_only(ctx->main.dev, n, 7, sink, grid, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&a, ctx->ev0, ctx->ev1) != hipSuccess ||
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_gather_raw(ctx->main.cells, ctx->main.hdr.m_cells, 3 * n, 11, policy, sink, grid, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&b, ctx->ev0, ctx->ev1) != hipSuccess) {
    ctx->err = "gather measurement failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(sink);
  *probe_ms = a;
  *raw_ms = b;
  return rc;
}

extern "C" int fc5_measure_copy(fc5_ctx* ctx, uint64_t bytes, double* ms) {
  if (!ctx || !ms || bytes < 8) return FC5_E_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const u64 n = bytes / 8;
  u64 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, n * 8) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "copy src");
  if (hipMalloc(&dst, n * 8) != hipSuccess) {
    (void)hipFree(src);
    FAIL(ctx, FC5_E_NOMEM, "copy dst");
  }
  float t = 0;
  int rc = FC5_OK;
  if (launch_fill_u64(src, n, 0x0123456789abcdefULL, ctx->stream) != 0 ||
      launch_copy_u64(src, dst, n, ctx->num_cu * 8, ctx->stream) != 0 ||  // warm-up
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_copy_u64(src, dst, n, ctx->num_cu * 8, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&t, ctx->ev0, ctx->ev1) != hipSuccess) {
    ctx->err = "copy measurement failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(src);
  (void)hipFree(dst);
  *ms = t;
  return rc;
}
