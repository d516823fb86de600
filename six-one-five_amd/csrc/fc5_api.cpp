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
#include "fc5_ctx.hpp"
#include "fc5_strict.hpp"

using namespace fc5;

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

extern "C" int fc5_ctx_create(int device, fc5_ctx** out) try {
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
      hipMalloc(&ctx->d_ctr, 16 * sizeof(Counters)) != hipSuccess ||
      hipMalloc(&ctx->d_p6, 751 * sizeof(u16)) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_ctr, 16 * sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_nhits, 8, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&ctx->d_nhits, 8) != hipSuccess || hipMalloc(&ctx->d_item_ctr, 8) != hipSuccess)
    return FC5_E_NOMEM;
  std::vector<u16> p6(751);
  for (u32 x = 0; x < 504; ++x) p6[x] = (u16)pow6_mod(x, 504);
  for (u32 x = 0; x < 247; ++x) p6[504 + x] = (u16)pow6_mod(x, 247);
  if (hipMemcpy(ctx->d_p6, p6.data(), p6.size() * sizeof(u16), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ctx->d_gate, bitmap.data(), kGateWordsU * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ctx->d_gate_ones, ones.data(), kGateWordsU * 4, hipMemcpyHostToDevice) != hipSuccess)
    return FC5_E_HIP;
  if (upload_sieve_tables() != 0) return FC5_E_HIP;
  (void)noalloc_setup(device);  // its failure is reported by fc5_run_units_noalloc alone
  *out = ctx.release();
  return FC5_OK;
} catch (...) {
  return abi_exception(nullptr);
}

extern "C" void fc5_ctx_destroy(fc5_ctx* ctx) {
  if (!ctx) return;
  serve_shutdown(ctx);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_part(ctx);
  free_table(ctx->main);
  free_table(ctx->side);
  for (void* p : {(void*)ctx->d_gate, (void*)ctx->d_gate_ones, (void*)ctx->d_ndg, (void*)ctx->d_hits,
                  (void*)ctx->d_ctr, (void*)ctx->d_nhits, (void*)ctx->d_item_ctr, (void*)ctx->d_p6})
    if (p) (void)hipFree(p);
  if (ctx->h_ctr) (void)hipHostFree(ctx->h_ctr);
  if (ctx->h_nhits) (void)hipHostFree(ctx->h_nhits);
  free_twin(ctx);
  free_exact(ctx);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* fc5_last_error(const fc5_ctx* ctx) {
  if (!ctx) return "null context";
  thread_local std::string text;  // valid until this thread's next call
  text = ctx->err.get();
  return text.c_str();
}

extern "C" int fc5_set_option(fc5_ctx* ctx, const char* key, int64_t v) try {
  if (!ctx || !key) return FC5_E_ARG;
  NOT_SERVING(ctx);
  const std::string k = key;
  if (k == "hit_cap") {
    if (v < 1 || v > (int64_t)1 << 32) FAIL(ctx, FC5_E_ARG, "hit_cap out of range");
    ctx->hit_cap = (u64)v;
  } else if (k == "use_gate") {
    ctx->use_gate = v ? 1 : 0;
  } else if (k == "strict_gate") {
    ctx->strict_gate = v ? 1 : 0;
  } else if (k == "sweep") {
    if (v != 0 && v != 1) FAIL(ctx, FC5_E_ARG, "sweep must be 0 (per-point bitmap test) or 1 (bit sieve)");
    ctx->sweep = (int)v;
  } else if (k == "verify") {
    ctx->verify = v ? 1 : 0;
  } else if (k == "verify_table") {
    ctx->verify_table = v ? 1 : 0;
  } else if (k == "stop_first") {
    ctx->stop_first = v ? 1 : 0;
  } else if (k == "load_policy") {
    if (v < -1 || v > 3) FAIL(ctx, FC5_E_ARG, "load_policy must be -1..3");
    ctx->load_policy = (int)v;
    ctx->main.dev.policy = table_policy(ctx->load_policy, ctx->main.hdr.m_cells);
    ctx->side.dev.policy = table_policy(ctx->load_policy, ctx->side.hdr.m_cells);
  } else if (k == "rows_per_item") {
    if (v != 0 && v != 4 && v != 8 && v != 16 && v != 32)
      FAIL(ctx, FC5_E_ARG, "rows_per_item must be 0 (by band size), 4, 8, 16 or 32");
    ctx->rows_per_item = (int)v;
  } else if (k == "use_graph") {
    ctx->use_graph = v ? 1 : 0;
  } else if (k == "stream_load_min") {
    if (v < 0) FAIL(ctx, FC5_E_ARG, "stream_load_min must be >= 0");
    ctx->stream_load_min = (u64)v;
  } else if (k == "twin_tri_gate") {
    ctx->twin_tri = v ? 1 : 0;
  } else if (k == "twin_plain") {
    ctx->twin_plain = v ? 1 : 0;
  } else if (k == "twin_ref_tiles") {
    if (v < 0 || v > 65535) FAIL(ctx, FC5_E_ARG, "twin_ref_tiles must be 0 (off) or a tile width 1..65535");
    ctx->twin_tile = (int)v;
  } else if (k == "twin_c4_div") {
    if (v != 3 && v != 4) FAIL(ctx, FC5_E_ARG, "twin_c4_div must be 3 or 4");
    ctx->twin_c4_div = (int)v;
  } else if (k == "probe_mode") {
    if (v < -1 || v > 1) FAIL(ctx, FC5_E_ARG, "probe_mode must be -1 (by table), 0 (gathers) or 1 (partitioned)");
    ctx->probe_mode = (int)v;
  } else if (k == "part_records") {
    if (v < (int64_t)kEmitChunk || v > (int64_t)0xFFFFFFFF) FAIL(ctx, FC5_E_ARG, "part_records must be 4096..2^32-1");
    if ((u64)v != ctx->part_records) free_part(ctx);  // the buffers are made for one capacity
    ctx->part_records = (u64)v;
  } else if (k == "part_hbm_limit") {
    if (v < 0) FAIL(ctx, FC5_E_ARG, "part_hbm_limit must be >= 0 (bytes; 0: no limit)");
    if ((u64)v != ctx->part_hbm_limit) free_part(ctx);
    ctx->part_hbm_limit = (u64)v;
  } else if (k == "blocks_per_cu") {
    if (v < 0 || v > 64) FAIL(ctx, FC5_E_ARG, "blocks_per_cu out of range");
    ctx->blocks_per_cu = (int)v;
  } else {
    FAIL(ctx, FC5_E_ARG, "unknown option " + k);
  }
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
static int upload_table(fc5_ctx* ctx, const StoreHeader& h, const u8* packed, size_t len,
                        DeviceTable& dst) {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  free_table(dst);
  dst.hdr = h;
  if (h.m_cells > (1ull << 40)) FAIL(ctx, FC5_E_FORMAT, "cell count out of range");  // 8.8 TB of cells
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
  dst.dev.policy = table_policy(ctx->load_policy, h.m_cells);
  dst.present = true;
  return FC5_OK;
}

static constexpr size_t kKeepHostTable = (size_t)4 << 30;

// The file's packed cells straight to the device: read in chunks into two
// pinned buffers, each copied up while the next is read, then widened into
// dst.  No host copy of the table is made (fc5_table_load of tables above
// kKeepHostTable; the reference reads the whole file into host memory first,
// fourcore_xor_store.hpp:147-181).
static int stream_table(fc5_ctx* ctx, std::FILE* f, const StoreHeader& h, DeviceTable& dst) {
  constexpr size_t kChunk = (size_t)64 << 20;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  dst = DeviceTable{};
  dst.hdr = h;
  if (h.m_cells > (1ull << 40)) FAIL(ctx, FC5_E_FORMAT, "cell count out of range");
  const size_t want = packed_bytes(h.m_cells, h.r);
  if (h.n_keys == 0 || h.m_cells < 3) {
    dst.present = true;
    return FC5_OK;
  }
  // +16 bytes of zero pad for 3-word reads, rounded to whole 8-byte words for the fill below
  const size_t words = ((want + 3) / 4 + 4 + 1) / 2 * 2;
  u8* staging = nullptr;
  u8* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool recorded[2] = {false, false};
  auto release = [&] {
    if (staging) (void)hipFree(staging);
    for (int k = 0; k < 2; ++k) {
      if (pin[k]) (void)hipHostFree(pin[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
  };
  if (hipMalloc(&staging, words * 4) != hipSuccess || hipMalloc(&dst.cells, h.m_cells * 8) != hipSuccess ||
      hipHostMalloc((void**)&pin[0], kChunk, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&pin[1], kChunk, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
    release();
    free_table(dst);
    FAIL(ctx, FC5_E_NOMEM, "device table or staging allocation");
  }
  int rc = FC5_OK;
  // the zero pad past the packed bytes (a fill kernel over the last 16-byte words)
  const size_t pad_from = want / 8 * 8;
  if (launch_fill_u64(reinterpret_cast<u64*>(staging + pad_from), (words * 4 - pad_from) / 8, 0, ctx->stream) != 0)
    rc = FC5_E_HIP;
  size_t off = 0;
  for (int k = 0; off < want && rc == FC5_OK; k ^= 1) {
    if (recorded[k] && hipEventSynchronize(ev[k]) != hipSuccess) {  // buffer k's last copy is done
      rc = FC5_E_HIP;
      break;
    }
    const size_t n = std::min(kChunk, want - off);
    if (std::fread(pin[k], 1, n, f) != n) {
      rc = FC5_E_IO;
      break;
    }
    if (hipMemcpyAsync(staging + off, pin[k], n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipEventRecord(ev[k], ctx->stream) != hipSuccess) {
      rc = FC5_E_HIP;
      break;
    }
    recorded[k] = true;
    off += n;
  }
  if (rc == FC5_OK &&
      (launch_expand_cells(reinterpret_cast<const u32*>(staging), h.m_cells, h.r, dst.cells, ctx->stream) != 0 ||
       hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = FC5_E_HIP;
  (void)hipStreamSynchronize(ctx->stream);
  release();
  if (rc != FC5_OK) {
    free_table(dst);
    ctx->err = rc == FC5_E_IO ? "short read of the table file" : "table upload failed";
    return rc;
  }
  dst.dev.cells = dst.cells;
  dst.dev.block = h.m_cells / 3;
  dst.dev.mask = fp_mask(h.r);
  dst.dev.seed = h.mix_seed;
  dst.dev.md = make_moddiv(dst.dev.block);
  dst.dev.present = dst.dev.block ? 1u : 0u;
  dst.dev.policy = table_policy(ctx->load_policy, h.m_cells);
  dst.present = true;
  return FC5_OK;
}

static int adopt_main(fc5_ctx* ctx, XorTable&& t) {
  const int rc = upload_table(ctx, t.hdr, t.packed.data(), t.packed.size(), ctx->main);
  if (rc != FC5_OK) return rc;
  ctx->host_main_valid = t.packed.size() <= kKeepHostTable;
  if (ctx->host_main_valid) ctx->host_main = std::move(t);
  else ctx->host_main = XorTable{};
  return FC5_OK;
}

static int fpr_device(fc5_ctx* ctx, const XorDev& x, int trials, int& hits);

extern "C" int fc5_table_load(fc5_ctx* ctx, const char* path, int expected_N) try {
  if (!ctx || !path) return FC5_E_ARG;
  NOT_SERVING(ctx);
  StoreHeader h;
  std::FILE* f = nullptr;
  std::string msg;
  const int st = open_bin(path, expected_N, h, f, msg);
  switch (st) {
    case kLoadOk: break;
    case kLoadOpen:
    case kLoadShort: FAIL(ctx, FC5_E_IO, msg);
    case kLoadNMismatch: FAIL(ctx, FC5_E_NMISMATCH, msg);
    default: FAIL(ctx, FC5_E_FORMAT, msg);
  }
  const size_t want = packed_bytes(h.m_cells, h.r);
  if (want <= ctx->stream_load_min) {  // small tables: read whole, smoke on the host, keep a host copy
    XorTable t;
    t.hdr = h;
    t.packed.assign(want, 0);
    const bool ok = !want || std::fread(t.packed.data(), 1, want, f) == want;
    std::fclose(f);
    if (!ok) FAIL(ctx, FC5_E_IO, "short read " + std::string(path));
    const int fpr = fpr_smoke(t);
    if (fpr < 0 || fpr > 16)
      FAIL(ctx, FC5_E_FPR, "xor FPR smoke too high (" + std::to_string(fpr) + ") - refuse to search");
    return adopt_main(ctx, std::move(t));
  }
  // large tables: streamed to the device, the same FPR smoke run there on the
  // new table before it replaces the current one (which stays on any failure)
  DeviceTable nt;
  int rc = stream_table(ctx, f, h, nt);
  std::fclose(f);
  if (rc != FC5_OK) return rc;
  int fpr = 0;
  rc = fpr_device(ctx, nt.dev, 200000, fpr);
  if (rc != FC5_OK) {
    free_table(nt);
    return rc;
  }
  if (fpr > 16) {
    free_table(nt);
    FAIL(ctx, FC5_E_FPR, "xor FPR smoke too high (" + std::to_string(fpr) + ") - refuse to search");
  }
  free_part(ctx);
  free_table(ctx->main);
  ctx->main = nt;
  ctx->host_main_valid = false;
  ctx->host_main = XorTable{};
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_table_upload(fc5_ctx* ctx, const fc5_store_header* hdr, const uint8_t* packed,
                                size_t len) try {
  if (!ctx || !hdr || (!packed && len)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  XorTable t;
  std::memcpy(&t.hdr, hdr, sizeof t.hdr);
  if (t.hdr.magic != 0x53353136u || t.hdr.version != 2 || t.hdr.kind != 1 || t.hdr.r == 0 || t.hdr.r > 64)
    FAIL(ctx, FC5_E_FORMAT, "bad header");
  t.packed.assign(packed, packed + len);
  return adopt_main(ctx, std::move(t));
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_table_build(fc5_ctx* ctx, int N, int r, const uint64_t* extra, size_t n_extra,
                               int on_gpu) try {
  if (!ctx || N < 1 || r < 1 || r > 64 || (!extra && n_extra)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (on_gpu) {
    HIPCHK(ctx, hipSetDevice(ctx->device));
    free_part(ctx);
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
    t.dev.policy = table_policy(ctx->load_policy, g.hdr.m_cells);
    t.present = true;
      return FC5_OK;
  }
  std::vector<u64> ex(extra, extra + n_extra);
  XorTable t;
  std::string msg;
  if (!build_pairs_cpu(N, (u32)r, ex, t, msg)) FAIL(ctx, FC5_E_BUILD, msg);
  return adopt_main(ctx, std::move(t));
} catch (...) {
  return abi_exception(ctx);
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

// A device table with no host copy to its file: packed on the device, then
// copied down in 64 MiB chunks through two pinned buffers, each written to the
// file while the next is copied (xor_save_file, fourcore_xor_store.hpp:126-143;
// the same bytes as table_packed + save_bin without a host copy of the table).
static int stream_save(fc5_ctx* ctx, const char* path) {
  constexpr size_t kChunk = (size_t)64 << 20;
  const StoreHeader& h = ctx->main.hdr;
  const u64 m = h.m_cells;
  const size_t want = packed_bytes(m, h.r);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t room = std::max<size_t>(want, (size_t)((m + 7) / 8) * h.r);
  u8* d = nullptr;
  u8* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::FILE* f = nullptr;
  auto release = [&] {
    (void)hipStreamSynchronize(ctx->stream);
    if (d) (void)hipFree(d);
    for (int k = 0; k < 2; ++k) {
      if (pin[k]) (void)hipHostFree(pin[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
    if (f) std::fclose(f);
  };
  if (hipMalloc(&d, room) != hipSuccess || hipHostMalloc((void**)&pin[0], kChunk, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&pin[1], kChunk, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
    release();
    FAIL(ctx, FC5_E_NOMEM, "pack buffers");
  }
  if (hipMemsetAsync(d, 0, room, ctx->stream) != hipSuccess ||
      gpu_pack_cells(ctx->main.cells, m, h.r, d, ctx->stream) != 0) {
    release();
    FAIL(ctx, FC5_E_HIP, "table pack failed");
  }
  f = std::fopen(path, "wb");
  if (!f || std::fwrite(&h, sizeof h, 1, f) != 1) {
    release();
    FAIL(ctx, FC5_E_IO, std::string("cannot write ") + path);
  }
  const size_t n_chunks = (want + kChunk - 1) / kChunk;
  auto enqueue = [&](size_t i) {
    const size_t off = i * kChunk, n = std::min(kChunk, want - off);
    return hipMemcpyAsync(pin[i % 2], d + off, n, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
           hipEventRecord(ev[i % 2], ctx->stream) == hipSuccess;
  };
  int rc = FC5_OK;
  if (n_chunks && !enqueue(0)) rc = FC5_E_HIP;
  for (size_t i = 0; i < n_chunks && rc == FC5_OK; ++i) {
    if (i + 1 < n_chunks && !enqueue(i + 1)) {  // into the other buffer, written out in the last step
      rc = FC5_E_HIP;
      break;
    }
    const size_t n = std::min(kChunk, want - i * kChunk);
    if (hipEventSynchronize(ev[i % 2]) != hipSuccess) rc = FC5_E_HIP;
    else if (std::fwrite(pin[i % 2], 1, n, f) != n) rc = FC5_E_IO;
  }
  std::FILE* g = f;
  f = nullptr;
  if (std::fclose(g) != 0 && rc == FC5_OK) rc = FC5_E_IO;  // buffered bytes that could not be written
  release();
  if (rc == FC5_E_IO) FAIL(ctx, rc, std::string("short write ") + path);
  if (rc != FC5_OK) FAIL(ctx, rc, "table copy-down failed");
  return FC5_OK;
}

extern "C" int fc5_table_save(fc5_ctx* ctx, const char* path) try {
  if (!ctx || !path) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  if (!ctx->host_main_valid && ctx->main.cells) return stream_save(ctx, path);
  XorTable t;
  const int rc = table_packed(ctx, t, true);
  if (rc != FC5_OK) return rc;
  std::string msg;
  if (!save_bin(path, t, msg)) FAIL(ctx, FC5_E_IO, msg);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_table_download(fc5_ctx* ctx, fc5_store_header* hdr, uint8_t* packed, size_t cap, size_t* len) try {
  if (!ctx || !len) return FC5_E_ARG;
  NOT_SERVING(ctx);
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
} catch (...) {
  return abi_exception(ctx);
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

extern "C" int fc5_table_query(fc5_ctx* ctx, const uint64_t* keys, size_t n, uint8_t* out) try {
  if (!ctx || (!keys && n) || (!out && n)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return device_query(ctx, ctx->main.dev, keys, n, out);
} catch (...) {
  return abi_exception(ctx);
}

// the FPR smoke of xor_fpr_smoke (fourcore_xor_store.hpp:54-65) on a device table
static int fpr_device(fc5_ctx* ctx, const XorDev& x, int trials, int& hits) {
  std::vector<u64> keys((size_t)trials);
  u64 v = 0x123456789abcdef0ULL;
  for (auto& k : keys) { v = v * kPhi + 1; k = v; }
  std::vector<u8> out((size_t)trials);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int rc = device_query(ctx, x, keys.data(), keys.size(), out.data());
  if (rc != FC5_OK) return rc;
  hits = 0;
  for (u8 b : out) hits += b;
  return FC5_OK;
}

extern "C" int fc5_table_fpr_smoke(fc5_ctx* ctx, int trials, int* hits) try {
  if (!ctx || !hits || trials < 0) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no table");
  return fpr_device(ctx, ctx->main.dev, trials, *hits);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_side_build(fc5_ctx* ctx, const uint64_t* keys, size_t n, int r) try {
  if (!ctx || (!keys && n) || r < 1 || r > 64) return FC5_E_ARG;
  NOT_SERVING(ctx);
  std::vector<u64> k(keys, keys + n);
  XorTable t;
  if (!build_xor_cpu(k, (u32)r, 0x51DE51DE51DEULL, 64, t)) FAIL(ctx, FC5_E_BUILD, "side table build stalled");
  return upload_table(ctx, t.hdr, t.packed.data(), t.packed.size(), ctx->side);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_side_clear(fc5_ctx* ctx) try {
  if (!ctx) return FC5_E_ARG;
  NOT_SERVING(ctx);
  free_table(ctx->side);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

// ---------------------------------------------------------------------------
// bands
// ---------------------------------------------------------------------------
int fc5::ensure_ndg(fc5_ctx* ctx, u32 dmax) {
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

int fc5::ensure_hits(fc5_ctx* ctx) {
  if (ctx->d_hits && ctx->d_hits_cap == ctx->hit_cap) return FC5_OK;
  if (ctx->d_hits) (void)hipFree(ctx->d_hits);
  ctx->d_hits = nullptr;
  if (hipMalloc(&ctx->d_hits, ctx->hit_cap * sizeof(HitRec)) != hipSuccess)
    FAIL(ctx, FC5_E_NOMEM, "hit buffer");
  ctx->d_hits_cap = ctx->hit_cap;
  return FC5_OK;
}

extern "C" int fc5_band_destroy(fc5_ctx* ctx, fc5_band* band) try {
  if (!band) return FC5_OK;
  if (ctx) (void)hipSetDevice(ctx->device);
  if (band->graph_exec) (void)hipGraphExecDestroy(band->graph_exec);
  if (band->graph) (void)hipGraphDestroy(band->graph);
  for (void* p : {(void*)band->d_units, (void*)band->d_plain, (void*)band->d_units3, (void*)band->d_segs})
    if (p) (void)hipFree(p);
  delete band;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

DevUnit5 fc5::dev_unit5(const Cls5Unit& U, u32 uid, u32 band) {
  DevUnit5 D{};
  D.Q_lo = (u64)U.Q;
  D.Q_hi = (u64)(U.Q >> 64);
  D.fmax1 = U.fmax1;
  D.fmax2 = U.fmax2;
  D.rows = 6u * (U.fmax1 / 729u + 1u);
  D.te = U.te;
  D.td = U.td;
  D.uid = uid;
  D.band = band;
  for (int k = 0; k < 6; ++k) D.res1[k] = (u16)U.res1[k];
  for (int k = 0; k < 4; ++k) D.res2[k] = (u8)U.res2[k];
  return D;
}

DevUnit3 fc5::dev_unit3(const Cls234Unit& U, u32 uid, u32 band) {
  DevUnit3 D{};
  D.Q_lo = (u64)U.Q;
  D.Q_hi = (u64)(U.Q >> 64);
  D.inv_lo = (u64)U.inv;
  D.inv_hi = (u64)(U.inv >> 64);
  D.fmax = U.fmax;
  D.lim = U.lim;
  D.rows = U.nres * (U.fmax / U.mod + 1u);
  D.mod = U.mod;
  D.dr = U.dr;
  D.shift = U.shift;
  D.nres = U.nres;
  D.uid = uid;
  D.band = band;
  for (u32 k = 0; k < U.nres; ++k) D.res[k] = (u16)U.res[k];
  return D;
}

int fc5::run_pair_jobs(fc5_ctx* ctx, const std::vector<PairJob>& jobs, const PairLaunch& launch,
                       std::vector<PairHit>& out, hipStream_t st) {
  if (!st) st = ctx->stream;
  out.clear();
  if (jobs.empty()) return FC5_OK;
  if (jobs.size() > 0xffffffffull) FAIL(ctx, FC5_E_ARG, "too many hits to verify");
  u32 cap = (u32)std::min<size_t>(std::max<size_t>(1024, 4 * jobs.size()), 1u << 30);
  PairJob* d_jobs = nullptr;
  PairHit* d_out = nullptr;
  u64* d_count = nullptr;
  auto release = [&] {
    for (void* p : {(void*)d_jobs, (void*)d_out, (void*)d_count})
      if (p) (void)hipFree(p);
    d_jobs = nullptr;
    d_out = nullptr;
    d_count = nullptr;
  };
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (hipMalloc(&d_jobs, jobs.size() * sizeof(PairJob)) != hipSuccess ||
        hipMalloc(&d_out, (size_t)cap * sizeof(PairHit)) != hipSuccess || hipMalloc(&d_count, 8) != hipSuccess) {
      release();
      FAIL(ctx, FC5_E_NOMEM, "pair search buffers");
    }
    u64 count = 0;
    if (hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice) != hipSuccess ||
        launch_fill_u64(d_count, 1, 0, st) != 0 ||
        launch(d_jobs, (u32)jobs.size(), d_out, cap, reinterpret_cast<u32*>(d_count), st) != 0 ||
        hipMemcpyAsync(&count, d_count, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      release();
      FAIL(ctx, FC5_E_HIP, "pair search");
    }
    if (count > cap) {  // more pairs than room: once more with room for all
      release();
      cap = (u32)count;
      continue;
    }
    out.resize((size_t)count);
    const hipError_t e =
        count ? hipMemcpy(out.data(), d_out, count * sizeof(PairHit), hipMemcpyDeviceToHost) : hipSuccess;
    release();
    if (e != hipSuccess) FAIL(ctx, FC5_E_HIP, "pair search read-back");
    return FC5_OK;
  }
  FAIL(ctx, FC5_E_HIP, "pair search overflowed twice");
}

int fc5::device_pairs(fc5_ctx* ctx, const std::vector<PairTarget>& targets,
                      std::vector<std::vector<std::pair<u32, u32>>>& pairs, std::vector<u8>* extra, hipStream_t st) {
  if (ctx->verify_table) return exact_table_pairs(ctx, targets, pairs, extra, st);
  if (extra) extra->assign(targets.size(), 0);
  pairs.assign(targets.size(), {});
  std::vector<PairJob> jobs;
  std::vector<size_t> which;
  for (size_t i = 0; i < targets.size(); ++i)
    if (targets[i].ok) {
      jobs.push_back(PairJob{(u64)targets[i].T, (u64)(targets[i].T >> 64), targets[i].hi, 0});
      which.push_back(i);
    }
  std::vector<PairHit> out;
  const int rc = run_pair_jobs(ctx, jobs, launch_pair_search, out, st);
  if (rc != FC5_OK) return rc;
  for (const PairHit& h : out) pairs[which[h.job]].emplace_back(h.a, h.b);
  for (auto& v : pairs) std::sort(v.begin(), v.end());
  return FC5_OK;
}

extern "C" int fc5_pair_search(fc5_ctx* ctx, const uint64_t* targets, const uint32_t* hi, size_t n, uint32_t* out,
                               size_t cap, size_t* n_out) try {
  if (!ctx || !n_out || (n && (!targets || !hi)) || (cap && !out)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<PairTarget> t(n);
  for (size_t i = 0; i < n; ++i) {
    // the device's sixth roots stay exact below 2^127 (every hit's target is below 2^118)
    if (targets[2 * i + 1] >> 63) FAIL(ctx, FC5_E_ARG, "pair search targets must be below 2^127");
    t[i] = PairTarget{((u128)targets[2 * i + 1] << 64) | targets[2 * i], hi[i], true};
  }
  std::vector<std::vector<std::pair<u32, u32>>> pairs;
  const int rc = device_pairs(ctx, t, pairs);
  if (rc != FC5_OK) return rc;
  size_t k = 0;
  for (size_t i = 0; i < n; ++i)
    for (auto [a, b] : pairs[i]) {
      if (k < cap) {
        out[3 * k] = (uint32_t)i;
        out[3 * k + 1] = a;
        out[3 * k + 2] = b;
      }
      ++k;
    }
  *n_out = k;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

// the band's launch segments: its class-5 items, then its class 2-4 items
static int upload_segments(fc5_ctx* ctx, fc5_band* b) {
  Segment seg[2];
  u32 n = 0;
  if (b->n_items) seg[n++] = Segment{0, 0, 0, b->n_dev, 0, 0};
  if (b->n_items3) seg[n++] = Segment{b->n_items, 0, 0, b->n_dev3, 1, 0};
  if (!b->d_segs && hipMalloc(&b->d_segs, 2 * sizeof(Segment)) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "segments");
  if (n) HIPCHK(ctx, hipMemcpy(b->d_segs, seg, n * sizeof(Segment), hipMemcpyHostToDevice));
  b->n_segs = n;
  return FC5_OK;
}

bool fc5::number_items(std::vector<DevUnit5>& dev, std::vector<DevUnit3>& dev3, int E, u32& n5, u32& n3) {
  u64 t5 = 0, t3 = 0;
  for (DevUnit5& D : dev) {
    D.item_base = (u32)t5;
    t5 += (D.rows + (u32)E - 1) / (u32)E;
    if (t5 > kMaxItems) return false;
  }
  for (DevUnit3& D : dev3) {
    D.item_base = (u32)t3;
    t3 += (D.rows + (u32)E - 1) / (u32)E;
    if (t3 > kMaxItems) return false;
  }
  if (t5 + t3 > kMaxItems) return false;
  n5 = (u32)t5;
  n3 = (u32)t3;
  return true;
}

int fc5::auto_rows_per_item(int num_cu, u64 rows, bool strict, bool sieve) {
  for (int E : {32, 16, 8}) {
    const u64 waves = (u64)num_cu * (u64)std::max(1, batched_blocks_per_cu(E, strict, sieve)) * 4u;
    if (rows / (u64)E >= 8 * waves) return E;
  }
  return 4;
}

// rows per work item for a band: the context option, or (option 0) by the band's size
static int band_rows_per_item(const fc5_ctx* ctx, const std::vector<DevUnit5>& dev,
                              const std::vector<DevUnit3>& dev3) {
  if (ctx->rows_per_item > 0) return ctx->rows_per_item;
  u64 rows = 0;
  for (const DevUnit5& D : dev) rows += D.rows;
  for (const DevUnit3& D : dev3) rows += D.rows;
  return auto_rows_per_item(ctx->num_cu, rows, ctx->strict_gate != 0, ctx->sweep != 0);
}

static int renumber(fc5_ctx* ctx, fc5_band* b) {
  const int E = band_rows_per_item(ctx, b->dev_h, b->dev3_h);
  if (b->rows == E) return FC5_OK;
  if (!number_items(b->dev_h, b->dev3_h, E, b->n_items, b->n_items3)) FAIL(ctx, FC5_E_ARG, kTooManyItems);
  if (!b->dev_h.empty())
    HIPCHK(ctx, hipMemcpy(b->d_units, b->dev_h.data(), b->dev_h.size() * sizeof(DevUnit5), hipMemcpyHostToDevice));
  if (!b->dev3_h.empty())
    HIPCHK(ctx, hipMemcpy(b->d_units3, b->dev3_h.data(), b->dev3_h.size() * sizeof(DevUnit3),
                          hipMemcpyHostToDevice));
  b->rows = E;
  return upload_segments(ctx, b);
}

static int band_create(fc5_ctx* ctx, const uint32_t* cls, const fc5_unit* units, size_t n, fc5_band** out) {
  if (!ctx || !out || (!units && n)) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (n > 0x7fffffffull) FAIL(ctx, FC5_E_ARG, "too many units in one band");
  *out = nullptr;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto b = std::make_unique<fc5_band>();
  b->n_in = n;
  b->in_units.assign(units, units + n);
  for (size_t i = 0; i < n; ++i) b->in_cls.push_back(cls ? cls[i] : 5u);
  b->cls_of.assign(n, 0);
  b->slot_of.assign(n, -1);
  std::vector<DevUnit5> dev;
  std::vector<DevUnit3> dev3;
  std::vector<PlainUnit> plain;
  std::vector<Cls5Unit> P5;
  std::vector<Cls234Unit> P3;
  std::vector<char> ok5, ok3;
  std::vector<int> slot;
  static_assert(sizeof(fc5_unit) == 16, "fc5_unit is two u64");
  pack_units(cls, reinterpret_cast<const u64*>(units), n, P5, ok5, P3, ok3, slot);
  for (size_t i = 0; i < n; ++i) {
    const int c = cls ? (int)cls[i] : 5;
    if (c == 5) {
      if (!ok5[(size_t)slot[i]]) { ++b->n_skipped; continue; }
      const Cls5Unit& U = P5[(size_t)slot[i]];
      b->cls_of[i] = 5;
      b->slot_of[i] = (int)b->units.size();
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
        dev.push_back(dev_unit5(U, (u32)i, 5));
      } else {
        b->irregular.push_back((u32)b->units.size());
      }
      b->units.push_back(U);
      b->uid.push_back((u32)i);
    } else if (c >= 2 && c <= 4) {
      if (!ok3[(size_t)slot[i]]) { ++b->n_skipped; continue; }
      const Cls234Unit& U = P3[(size_t)slot[i]];
      b->cls_of[i] = (u8)c;
      b->slot_of[i] = (int)b->units3.size();
      b->max_lim3 = std::max(b->max_lim3, U.lim);
      if (U.regular) {
        dev3.push_back(dev_unit3(U, (u32)i, (u32)c));
      } else {
        b->irregular3.push_back((u32)b->units3.size());
      }
      b->units3.push_back(U);
      b->uid3.push_back((u32)i);
    } else {
      ++b->n_skipped;
    }
  }
  b->n_dev = (u32)dev.size();
  b->n_plain = (u32)plain.size();
  b->n_dev3 = (u32)dev3.size();
  const int E = band_rows_per_item(ctx, dev, dev3);
  if (!number_items(dev, dev3, E, b->n_items, b->n_items3)) FAIL(ctx, FC5_E_ARG, kTooManyItems);
  b->rows = E;
  int rc = FC5_OK;
  auto upload = [&](void** dst, const void* src, size_t bytes) {
    if (!bytes || rc != FC5_OK) return;
    if (hipMalloc(dst, bytes) != hipSuccess) { ctx->err = "unit upload"; rc = FC5_E_NOMEM; return; }
    if (hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) { ctx->err = "unit copy"; rc = FC5_E_HIP; }
  };
  upload((void**)&b->d_units, dev.data(), dev.size() * sizeof(DevUnit5));
  upload((void**)&b->d_plain, plain.data(), plain.size() * sizeof(PlainUnit));
  upload((void**)&b->d_units3, dev3.data(), dev3.size() * sizeof(DevUnit3));
  b->dev_h = std::move(dev);
  b->dev3_h = std::move(dev3);
  if (rc == FC5_OK) rc = upload_segments(ctx, b.get());
  if (rc == FC5_OK && !b->units.empty()) rc = ensure_ndg(ctx, b->max_fmax2);
  if (rc != FC5_OK) {
    fc5_band_destroy(ctx, b.release());
    return rc;
  }
  *out = b.release();
  return FC5_OK;
}

extern "C" int fc5_band_create(fc5_ctx* ctx, const fc5_unit* units, size_t n, fc5_band** out) try {
  return band_create(ctx, nullptr, units, n, out);
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_band_create_mixed(fc5_ctx* ctx, const uint32_t* cls, const fc5_unit* units, size_t n,
                                     fc5_band** out) try {
  if (!cls && n) return FC5_E_ARG;
  return band_create(ctx, cls, units, n, out);
} catch (...) {
  return abi_exception(ctx);
}

static PlainArgs plain_args(fc5_ctx* ctx) {
  PlainArgs p{};
  p.use_gate = ctx->use_gate;
  p.strict = ctx->strict_gate;
  p.gate = ctx->gw;
  p.main = ctx->main.dev;
  p.side = ctx->side.present ? ctx->side.dev : XorDev{};
  p.inv216 = inv_mod_2_64(85766121ULL);
  const u64 m504 = 504ULL * kM42;
  p.c_m504 = (u64)(((u128)1 << 64) % m504);
  p.c_m247 = (u64)(((u128)1 << 64) % 247u);
  const u64 p6 = (u64)pow6_mod(42, 247);  // inverse of 42^6 mod 247
  u64 inv = 1;
  for (u64 x = 1; x < 247; ++x)
    if (p6 * x % 247 == 1) { inv = x; break; }
  p.inv42_247 = inv;
  p.hits = ctx->d_hits;
  p.hit_cap = ctx->hit_cap;
  p.nhits = ctx->d_nhits;
  p.ctr = ctx->d_ctr + 8 + 5;
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

// classes 2-4 in the reference shape: host expansion of every unit into jobs
// (fourcore_find_v4.cu:165-205), windows (:401-425), 256 candidates per
// synchronous k_find3 launch (:492, CHUNK :584).  failT is counted on the host.
static int run_plain_find3(fc5_ctx* ctx, fc5_band* b, const std::vector<u32>& which, u64 fail_by_cls[6],
                           u64 jobs_by_cls[6]) {
  constexpr int kChunk = 256;
  const bool all = which.empty();
  const size_t total = all ? b->units3.size() : which.size();
  if (!total) return FC5_OK;
  PlainCand3* h_c = nullptr;
  PlainCand3* d_c = nullptr;
  if (hipHostMalloc((void**)&h_c, kChunk * sizeof(PlainCand3), hipHostMallocDefault) != hipSuccess)
    FAIL(ctx, FC5_E_NOMEM, "pinned candidates");
  if (hipMalloc(&d_c, kChunk * sizeof(PlainCand3)) != hipSuccess) {
    (void)hipHostFree(h_c);
    FAIL(ctx, FC5_E_NOMEM, "candidates");
  }
  PlainF3Args a{};
  a.use_gate = ctx->use_gate;
  a.strict = ctx->strict_gate;
  a.gate = ctx->gw;
  a.p6 = ctx->d_p6;
  a.main = ctx->main.dev;
  a.side = ctx->side.present ? ctx->side.dev : XorDev{};
  a.hits = ctx->d_hits;
  a.hit_cap = ctx->hit_cap;
  a.nhits = ctx->d_nhits;
  a.ctr = ctx->d_ctr + 8;
  a.cands = d_c;
  int rc = FC5_OK, n = 0;
  auto flush = [&]() {
    if (!n || rc != FC5_OK) return;
    a.n = n;
    if (hipMemcpyAsync(d_c, h_c, n * sizeof(PlainCand3), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        launch_find3_plain(a, ctx->stream) != 0 || hipStreamSynchronize(ctx->stream) != hipSuccess) {
      ctx->err = "plain find3 launch";
      rc = FC5_E_HIP;
    }
    n = 0;
  };
  std::vector<std::pair<u32, u128>> jobs;
  for (size_t k = 0; k < total && rc == FC5_OK; ++k) {
    const Cls234Unit& U = b->units3[all ? k : which[k]];
    jobs.clear();
    u64 fail = 0;
    expand_cls234(U, jobs, fail);
    fail_by_cls[U.cls] += fail;
    jobs_by_cls[U.cls] += jobs.size();
    const u32 uid = b->uid3[all ? k : which[k]];
    for (const auto& [d, T] : jobs) {
      u32 lo, hi;
      find3_window(T, U.lim, lo, hi);
      if (lo > hi) continue;
      PlainCand3& C = h_c[n++];
      C.q_lo = (u64)T;
      C.lo = lo;
      C.hi = hi;
      C.r504 = (u32)(T % 504);
      C.r247 = (u32)(T % 247);
      C.uid = uid;
      C.d = d;
      C.band = (u32)U.cls;
      const StrictRow sr = strict_row((u64)(T >> 64), (u64)T);
      C.srow = sr.w0;
      C.srow1 = sr.w1;
      if (n == kChunk) flush();
    }
  }
  flush();
  (void)hipFree(d_c);
  (void)hipHostFree(h_c);
  return rc;
}

static BatchedArgs batched_args(fc5_ctx* ctx, fc5_band* b) {
  BatchedArgs a{};
  a.units = b->d_units;
  a.segs = b->d_segs;
  a.n_segs = b->n_segs;
  a.total = b->n_items + b->n_items3;
  a.gate = ctx->use_gate ? ctx->d_gate : ctx->d_gate_ones;
  a.ndg = ctx->d_ndg;
  a.ndg_K = ctx->ndg_K;
  a.use_gate = (u32)ctx->use_gate;
  a.strict = (u32)ctx->strict_gate;
  a.sieve = (u32)ctx->sweep;
  a.main = ctx->main.dev;
  a.side = ctx->side.present ? ctx->side.dev : XorDev{};
  a.hits = ctx->d_hits;
  a.hit_cap = ctx->hit_cap;
  a.nhits = ctx->d_nhits;
  a.ctr = ctx->d_ctr;
  a.item_ctr = ctx->d_item_ctr;
  a.inv729 = inv_mod_2_64(729);
  a.units3 = b->d_units3;
  return a;
}

// counters, hit slot and work-queue counter cleared by a fill kernel (no
// memset nodes, so the same sequence is valid inside a captured graph)
static int reset_counters(fc5_ctx* ctx) {
  if (launch_fill_u64((u64*)ctx->d_ctr, 16 * sizeof(Counters) / 8, 0, ctx->stream) != 0 ||
      launch_fill_u64(ctx->d_nhits, 1, 0, ctx->stream) != 0 ||
      launch_fill_u64((u64*)ctx->d_item_ctr, 1, 0, ctx->stream) != 0)
    FAIL(ctx, FC5_E_HIP, "counter reset");
  return FC5_OK;
}

// capture reset + persistent kernel + counter read-back as one graph per band;
// captured again whenever a baked argument (tables, options, grid) changes
static int ensure_graph(fc5_ctx* ctx, fc5_band* b, const BatchedArgs& a, int grid) {
  std::vector<u64> key(sizeof(BatchedArgs) / 8 + 2);
  std::memcpy(key.data(), &a, sizeof(BatchedArgs));
  key[key.size() - 2] = (u64)grid;
  key[key.size() - 1] = (u64)b->rows;
  if (b->graph_exec && key == b->graph_key) return FC5_OK;
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = reset_counters(ctx);
  if (rc == FC5_OK && launch_cls5_batched(a, grid, b->rows, ctx->stream) != 0) rc = FC5_E_HIP;
  if (rc == FC5_OK &&
      (hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, 16 * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream) !=
           hipSuccess ||
       hipMemcpyAsync(ctx->h_nhits, ctx->d_nhits, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess))
    rc = FC5_E_HIP;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != FC5_OK || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    FAIL(ctx, FC5_E_HIP, "graph capture failed");
  }
  if (hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    FAIL(ctx, FC5_E_HIP, "graph instantiate failed");
  }
  b->graph = g;
  b->graph_key = key;
  return FC5_OK;
}

// The batched kernel in the reference's launch pattern: one launch per 16
// units of one arena (fourcore_cls5_gpu_v4.cu:468-501), each followed by a
// synchronize.  Same kernel and counters as the persistent launch; it
// isolates what batching into one launch buys from what the T-space kernel does.
static int run_chunked(fc5_ctx* ctx, fc5_band* b, BatchedArgs a, int bpc) {
  constexpr u32 kChunk = 16;
  std::vector<Segment> segs;
  std::vector<u32> items;
  for (u32 u0 = 0; u0 < b->n_dev; u0 += kChunk) {
    const u32 n = std::min(kChunk, b->n_dev - u0);
    const u32 i0 = b->dev_h[u0].item_base, i1 = u0 + n < b->n_dev ? b->dev_h[u0 + n].item_base : b->n_items;
    segs.push_back(Segment{0, i0, u0, n, 0, 0});
    items.push_back(i1 - i0);
  }
  for (u32 u0 = 0; u0 < b->n_dev3; u0 += kChunk) {
    const u32 n = std::min(kChunk, b->n_dev3 - u0);
    const u32 i0 = b->dev3_h[u0].item_base, i1 = u0 + n < b->n_dev3 ? b->dev3_h[u0 + n].item_base : b->n_items3;
    segs.push_back(Segment{0, i0, u0, n, 1, 0});
    items.push_back(i1 - i0);
  }
  if (segs.empty()) return FC5_OK;
  Segment* d_segs = nullptr;
  if (hipMalloc(&d_segs, segs.size() * sizeof(Segment)) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "chunk segments");
  int rc = FC5_OK;
  if (hipMemcpy(d_segs, segs.data(), segs.size() * sizeof(Segment), hipMemcpyHostToDevice) != hipSuccess)
    rc = FC5_E_HIP;
  for (size_t i = 0; i < segs.size() && rc == FC5_OK; ++i) {
    if (!items[i]) continue;
    a.segs = d_segs + i;
    a.n_segs = 1;
    a.total = items[i];
    const int grid = std::max(1, std::min<int>(ctx->num_cu * bpc, (int)items[i]));
    if (launch_fill_u64((u64*)ctx->d_item_ctr, 1, 0, ctx->stream) != 0 ||
        launch_cls5_batched(a, grid, b->rows, ctx->stream) != 0 ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = FC5_E_HIP;
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_segs);
  if (rc != FC5_OK) ctx->err = "chunked launch";
  return rc;
}

extern "C" int fc5_band_run(fc5_ctx* ctx, fc5_band* b, int mode, fc5_result* res) try {
  if (!ctx || !b || !res) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (mode != FC5_MODE_PLAIN && mode != FC5_MODE_BATCHED && mode != FC5_MODE_CHUNKED)
    FAIL(ctx, FC5_E_ARG, "unknown mode");
  if (!ctx->main.present) FAIL(ctx, FC5_E_STATE, "no pair table loaded");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_hits(ctx);
  if (rc != FC5_OK) return rc;
  std::memset(res, 0, sizeof *res);
  std::memset(ctx->class_res, 0, sizeof ctx->class_res);
  ctx->hits.clear();
  ctx->sols.clear();
  u64 fail_host[6] = {0}, jobs_host[6] = {0};
  rc = renumber(ctx, b);
  if (rc != FC5_OK) return rc;
  const bool batched = mode != FC5_MODE_PLAIN;  // the T-space kernel, in one launch or in chunks
  const bool any_items = b->n_items + b->n_items3 > 0;
  BatchedArgs a = batched_args(ctx, b);
  const int bpc = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : batched_blocks_per_cu(b->rows, ctx->strict_gate != 0, ctx->sweep != 0);
  const int grid = std::max(1, std::min<int>(ctx->num_cu * bpc, (int)(b->n_items + b->n_items3)));
  const bool part = mode == FC5_MODE_BATCHED && any_items && use_partitioned(ctx);
  if (!part) {  // the partitioned-probing figures describe the last run only
    std::fill(std::begin(ctx->part_ms), std::end(ctx->part_ms), 0.0);
    std::fill(std::begin(ctx->part_stage_ms), std::end(ctx->part_stage_ms), 0.0);
    std::fill(std::begin(ctx->part_counts), std::end(ctx->part_counts), 0);
  }
  const bool graph = mode == FC5_MODE_BATCHED && any_items && ctx->use_graph && !part && b->irregular.empty() &&
                     b->irregular3.empty();
  HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (graph) {
    rc = ensure_graph(ctx, b, a, grid);
    if (rc != FC5_OK) return rc;
    HIPCHK(ctx, hipGraphLaunch(b->graph_exec, ctx->stream));
  } else {
    rc = reset_counters(ctx);
    if (rc != FC5_OK) return rc;
    if (batched) {
      if (mode == FC5_MODE_CHUNKED) {
        rc = run_chunked(ctx, b, a, bpc);
        if (rc != FC5_OK) return rc;
      } else if (part) {
        rc = run_partitioned(ctx, a, grid, b->rows);
        if (rc != FC5_OK) return rc;
      } else if (any_items && launch_cls5_batched(a, grid, b->rows, ctx->stream) != 0) {
        FAIL(ctx, FC5_E_HIP, "batched launch");
      }
      if (!b->irregular.empty()) {
        rc = run_plain_units(ctx, b, b->irregular, false);
        if (rc != FC5_OK) return rc;
      }
      if (!b->irregular3.empty()) {
        rc = run_plain_find3(ctx, b, b->irregular3, fail_host, jobs_host);
        if (rc != FC5_OK) return rc;
      }
    } else {
      if (b->n_plain) {
        rc = run_plain_units(ctx, b, {}, true);
        if (rc != FC5_OK) return rc;
      }
      rc = run_plain_find3(ctx, b, {}, fail_host, jobs_host);
      if (rc != FC5_OK) return rc;
    }
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_ctr, ctx->d_ctr, 16 * sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->h_nhits, ctx->d_nhits, 8, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  Counters c[16];
  std::memcpy(c, ctx->h_ctr, sizeof c);
  const u64 nh = *ctx->h_nhits;
  float ms = 0;
  HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const u64 stored = std::min<u64>(nh, ctx->hit_cap);
  std::vector<HitRec> raw(stored);
  if (stored)
    HIPCHK(ctx, hipMemcpy(raw.data(), ctx->d_hits, stored * sizeof(HitRec), hipMemcpyDeviceToHost));
  for (int k = 2; k <= 5; ++k) {
    const Counters& B = c[k];      // batched kernel: gated = positives - probes
    const Counters& P = c[8 + k];  // plain kernels: gated counted directly
    fc5_result& r = ctx->class_res[k];
    r.calls = B.calls + P.calls;
    r.neg = B.neg + P.neg;
    r.zero = B.zero + P.zero;
    r.probes = B.probes + P.probes;
    r.gated = (B.pos - B.probes) + P.gated;
    r.maybe = B.maybe + P.maybe;
    r.side_maybe = B.side_maybe + P.side_maybe;
    r.jobs = B.jobs + jobs_host[k];
    r.fail_t = B.failT + fail_host[k];
    r.strict_out = ctx->strict_gate ? (B.probes + P.probes) - (B.queries + P.queries) : 0;
    res->calls += r.calls;
    res->neg += r.neg;
    res->zero += r.zero;
    res->probes += r.probes;
    res->gated += r.gated;
    res->maybe += r.maybe;
    res->side_maybe += r.side_maybe;
    res->jobs += r.jobs;
    res->fail_t += r.fail_t;
    res->strict_out += r.strict_out;
  }
  for (size_t i = 0; i < b->n_in; ++i) {
    const int k = b->cls_of[i];
    if (k) ctx->class_res[k].units++;
  }
  res->n_hits = nh;
  res->hits_dropped = nh - stored;
  res->units = b->units.size() + b->units3.size();
  res->units_skipped = b->n_skipped;
  res->units_irregular = b->irregular.size() + b->irregular3.size();
  res->kernel_ms = ms;
  ctx->hits.resize(stored);
  for (u64 i = 0; i < stored; ++i)
    ctx->hits[i] = fc5_hit{raw[i].unit, raw[i].d, raw[i].e, raw[i].flags & 0xffu, raw[i].fp};
  std::sort(ctx->hits.begin(), ctx->hits.end(), [](const fc5_hit& x, const fc5_hit& y) {
    if (x.unit != y.unit) return x.unit < y.unit;
    if (x.e != y.e) return x.e < y.e;
    if (x.d != y.d) return x.d < y.d;
    return x.flags < y.flags;
  });
  if (ctx->verify && !ctx->hits.empty()) {
    const int N = ctx->main.hdr.N ? (int)ctx->main.hdr.N : (int)std::max(b->max_lim, b->max_lim3);
    std::vector<VerifyJob> jobs;
    std::vector<size_t> hit_of;
    for (size_t i = 0; i < ctx->hits.size(); ++i) {
      const fc5_hit& h = ctx->hits[i];
      if (!(h.flags & 1u) || h.unit >= b->n_in) continue;
      const int k = b->cls_of[h.unit], slot = b->slot_of[h.unit];
      if (k == 5) {
        const Cls5Unit& U = b->units[(size_t)slot];
        jobs.push_back(VerifyJob{5, U.B, U.u, U.lim, h.d, h.e, h.fp});
      } else if (k >= 2) {
        const Cls234Unit& U = b->units3[(size_t)slot];
        jobs.push_back(VerifyJob{k, U.B, U.u, U.lim, h.d, h.e, h.fp});
      } else {
        continue;
      }
      hit_of.push_back(i);
    }
    std::vector<PairTarget> targets;
    verify_targets(N, jobs, targets);
    std::vector<std::vector<std::pair<u32, u32>>> pairs;
    std::vector<u8> extra;
    rc = device_pairs(ctx, targets, pairs, &extra);
    if (rc != FC5_OK) return rc;
    for (size_t i = 0; i < extra.size(); ++i)
      if (extra[i]) ctx->hits[hit_of[i]].flags |= FC5_HIT_EXTRA;
    std::vector<Solution> sols;
    verify_finish_all(jobs, pairs, ctx->stop_first != 0, sols);
    for (const Solution& S : sols) {
      fc5_solution o{};
      o.B = S.B;
      o.u = S.u;
      o.cls = S.cls;
      o.d = S.d;
      o.e = S.e;
      o.a = S.a;
      o.b = S.b;
      o.c = S.c;
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
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_autotune(fc5_ctx* ctx, fc5_band* b, int sample_every, fc5_tune* out) try {
  if (!ctx || !b || !out || sample_every < 1) return FC5_E_ARG;
  NOT_SERVING(ctx);
  std::vector<fc5_unit> su;
  std::vector<u32> sc;
  for (size_t i = 0; i < b->n_in; i += (size_t)sample_every) {
    su.push_back(b->in_units[i]);
    sc.push_back(b->in_cls[i]);
  }
  fc5_band* sb = nullptr;
  int rc = band_create(ctx, sc.data(), su.data(), su.size(), &sb);
  if (rc != FC5_OK) return rc;
  const int keep_verify = ctx->verify, keep_rows = ctx->rows_per_item, keep_bpc = ctx->blocks_per_cu;
  const int auto_E = band_rows_per_item(ctx, b->dev_h, b->dev3_h);  // the shape used without tuning
  ctx->verify = 0;
  std::memset(out, 0, sizeof *out);
  out->best_ms = 1e30;
  fc5_result r;
  for (int E : {8, 4, 16, 32}) {
    const int occ = batched_blocks_per_cu(E, ctx->strict_gate != 0, ctx->sweep != 0);
    for (int bpc : {occ, std::max(1, occ / 2)}) {
      ctx->rows_per_item = E;
      ctx->blocks_per_cu = bpc;
      double ms = 1e30;
      for (int rep = 0; rep < 2 && rc == FC5_OK; ++rep) {
        rc = fc5_band_run(ctx, sb, FC5_MODE_BATCHED, &r);
        ms = std::min(ms, r.kernel_ms);
      }
      if (rc != FC5_OK) break;
      ++out->n_tried;
      if (E == auto_E && bpc == occ) out->default_ms = ms;
      if (ms < out->best_ms) {
        out->best_ms = ms;
        out->rows_per_item = E;
        out->blocks_per_cu = bpc;
      }
    }
    if (rc != FC5_OK) break;
  }
  ctx->verify = keep_verify;
  fc5_band_destroy(ctx, sb);
  if (rc != FC5_OK) {
    ctx->rows_per_item = keep_rows;
    ctx->blocks_per_cu = keep_bpc;
    return rc;
  }
  // a shape within 3% of the untuned one on the sample is no reliable gain on the whole
  // band (the sample's launches are short, so their tails weigh more): keep the untuned shape
  if (out->default_ms > 0 && out->best_ms > 0.97 * out->default_ms) {
    out->rows_per_item = auto_E;
    out->blocks_per_cu = batched_blocks_per_cu(auto_E, ctx->strict_gate != 0, ctx->sweep != 0);
    out->best_ms = out->default_ms;
    ctx->rows_per_item = keep_rows;
    ctx->blocks_per_cu = keep_bpc;
    return FC5_OK;
  }
  ctx->rows_per_item = out->rows_per_item;
  ctx->blocks_per_cu = out->blocks_per_cu;
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_last_class_result(fc5_ctx* ctx, int cls, fc5_result* out) try {
  if (!ctx || !out || cls < 2 || cls > 5) return FC5_E_ARG;
  *out = ctx->class_res[cls];
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_run_units(fc5_ctx* ctx, const fc5_unit* units, size_t n, int mode, fc5_result* res) try {
  fc5_band* b = nullptr;
  int rc = fc5_band_create(ctx, units, n, &b);
  if (rc != FC5_OK) return rc;
  rc = fc5_band_run(ctx, b, mode, res);
  fc5_band_destroy(ctx, b);
  return rc;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_last_hits(fc5_ctx* ctx, fc5_hit* out, size_t cap, size_t* n) try {
  if (!ctx || !n || (!out && cap)) return FC5_E_ARG;
  *n = ctx->hits.size();
  std::copy_n(ctx->hits.begin(), std::min(cap, ctx->hits.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_last_solutions(fc5_ctx* ctx, fc5_solution* out, size_t cap, size_t* n) try {
  if (!ctx || !n || (!out && cap)) return FC5_E_ARG;
  *n = ctx->sols.size();
  std::copy_n(ctx->sols.begin(), std::min(cap, ctx->sols.size()), out);
  return FC5_OK;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_units_calls(const fc5_unit* units, size_t n, uint64_t* calls) try {
  if ((!units && n) || !calls) return FC5_E_ARG;
  u64 c = 0;
  for (size_t i = 0; i < n; ++i) {
    Cls5Unit U;
    if (pack_cls5(units[i].B, units[i].u, U)) c += cls5_calls(U);
  }
  *calls = c;
  return FC5_OK;
} catch (...) {
  return abi_exception(nullptr);
}

extern "C" int fc5_emit_units(uint64_t lo, uint64_t hi, unsigned cls_mask, double lo_frac, double hi_frac,
                              uint32_t* out_cls, uint64_t* out_B, uint64_t* out_u, size_t cap, size_t* n) try {
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
} catch (...) {
  return abi_exception(nullptr);
}

extern "C" int fc5_build_pairs_file(int N, int r, const char* path, fc5_store_header* hdr_out) try {
  if (N < 1 || r < 1 || r > 64 || !path) return FC5_E_ARG;
  XorTable t;
  std::string msg;
  if (!build_pairs_cpu(N, (u32)r, {}, t, msg)) return FC5_E_BUILD;
  if (!save_bin(path, t, msg)) return FC5_E_IO;
  if (hdr_out) std::memcpy(hdr_out, &t.hdr, sizeof *hdr_out);
  return FC5_OK;
} catch (...) {
  return abi_exception(nullptr);
}

// ---------------------------------------------------------------------------
// measurement
// ---------------------------------------------------------------------------
extern "C" int fc5_measure_gathers(fc5_ctx* ctx, uint64_t n, double* probe_ms, double* raw_ms) try {
  if (!ctx || !probe_ms || !raw_ms || n == 0) return FC5_E_ARG;
  NOT_SERVING(ctx);
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
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_measure_gathers_region(fc5_ctx* ctx, uint64_t region_cells, uint64_t n, double* ms) try {
  if (!ctx || !ms || n == 0 || region_cells == 0) return FC5_E_ARG;
  NOT_SERVING(ctx);
  if (!ctx->main.present || !ctx->main.cells) FAIL(ctx, FC5_E_STATE, "no device table");
  if (region_cells > ctx->main.hdr.m_cells) FAIL(ctx, FC5_E_ARG, "region larger than the table");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned long long* sink = nullptr;
  if (hipMalloc(&sink, 8) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "sink");
  const int grid = ctx->num_cu * 8;
  const u32 policy = ctx->main.dev.policy;
  float t = 0;
  int rc = FC5_OK;
  if (launch_gather_raw(ctx->main.cells, region_cells, n / 4, 3, policy, sink, grid, ctx->stream) != 0 ||  // warm-up
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_gather_raw(ctx->main.cells, region_cells, n, 13, policy, sink, grid, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&t, ctx->ev0, ctx->ev1) != hipSuccess) {
    ctx->err = "gather measurement failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(sink);
  *ms = t;
  return rc;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_measure_read(fc5_ctx* ctx, uint64_t bytes, double* ms) try {
  if (!ctx || !ms || bytes < 16) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const u64 n = bytes / 16 * 2;  // u64 words of a whole number of 16-byte words
  u64* src = nullptr;
  unsigned long long* sink = nullptr;
  if (hipMalloc(&src, n * 8) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "read buffer");
  if (hipMalloc(&sink, 8) != hipSuccess) {
    (void)hipFree(src);
    FAIL(ctx, FC5_E_NOMEM, "sink");
  }
  const int grid = ctx->num_cu * 32;
  float t = 0;
  int rc = FC5_OK;
  if (launch_fill_u64(src, n, 0x0123456789abcdefULL, ctx->stream) != 0 ||
      launch_read_stream(src, n * 8, sink, grid, ctx->stream) != 0 ||  // warm-up
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_read_stream(src, n * 8, sink, grid, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&t, ctx->ev0, ctx->ev1) != hipSuccess) {
    ctx->err = "read measurement failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(src);
  (void)hipFree(sink);
  *ms = t;
  return rc;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_measure_valu(fc5_ctx* ctx, int kind, double* instr_per_s) try {
  if (!ctx || !instr_per_s || kind < 0 || kind > 3) return FC5_E_ARG;
  NOT_SERVING(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned long long* sink = nullptr;
  if (hipMalloc(&sink, 8) != hipSuccess) FAIL(ctx, FC5_E_NOMEM, "sink");
  const int grid = ctx->num_cu * 8;  // 32 waves per CU: the occupancy limit of 256-thread groups
  const u32 iters = (u32)(4500000u / kValuInstrs[kind]);
  float t = 0;
  int rc = FC5_OK;
  if (launch_valu(kind, iters / 8, 1, sink, grid, ctx->stream) != 0 ||  // warm-up
      hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
      launch_valu(kind, iters, 0x2545F491u, sink, grid, ctx->stream) != 0 ||
      hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->ev1) != hipSuccess ||
      hipEventElapsedTime(&t, ctx->ev0, ctx->ev1) != hipSuccess) {
    ctx->err = "VALU measurement failed";
    rc = FC5_E_HIP;
  }
  (void)hipFree(sink);
  const double waves = (double)grid * 256 / 64;
  *instr_per_s = t > 0 ? waves * iters * kValuInstrs[kind] / (t / 1000.0) : 0.0;
  return rc;
} catch (...) {
  return abi_exception(ctx);
}

extern "C" int fc5_measure_copy(fc5_ctx* ctx, uint64_t bytes, double* ms) try {
  if (!ctx || !ms || bytes < 8) return FC5_E_ARG;
  NOT_SERVING(ctx);
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
} catch (...) {
  return abi_exception(ctx);
}
