// xor_build_save -- build the packed xor pair table and write StoreHeader v2 +
// packed bytes.  Same flags, limits and output as the reference program
// (xor_build_save.cpp:14-57); the bytes are identical to the reference builder.
// Addition: --gpu [--device K] builds on the device (seconds instead of hours
// at N ~ 1e5): a valid filter over the same keys, with different bytes.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "fc5.h"
#include "fc5_host.hpp"

using namespace fc5;

static void usage() {
  std::fprintf(stderr,
               "usage: xor_build_save --N n [--r BITS] [--out FILE] [--gpu [--device K]]\n"
               "  N=3000000/42 → 71428 for B through 3.0M\n"
               "  N=2800000/42 → 66666 for B through 2.8M\n");
}

int main(int argc, char** argv) {
  int N = 0, r = 48, device = 0;
  bool gpu = false;
  std::string out = "runs/xor_packed.bin";
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* { return (i + 1 < argc) ? argv[++i] : ""; };
    if (a == "--N") N = std::atoi(next());
    else if (a == "--r") r = std::atoi(next());
    else if (a == "--out") out = next();
    else if (a == "--gpu") gpu = true;
    else if (a == "--device") device = std::atoi(next());
    else if (a == "-h" || a == "--help") { usage(); return 0; }
    else { std::fprintf(stderr, "unknown %s\n", a.c_str()); usage(); return 1; }
  }
  if (N < 4 || N > 120000) {
    std::fprintf(stderr, "need --N in 4..%d\n", 120000);
    usage();
    return 1;
  }
  if (r < 8 || r > 64) { std::fprintf(stderr, "--r must be 8..64\n"); return 1; }
  std::fprintf(stderr, "[xor_build_save] building N=%d r=%d → %s\n", N, r, out.c_str());
  const auto t0 = std::chrono::steady_clock::now();
  if (gpu) {
    fc5_ctx* ctx = nullptr;
    fc5_store_header h{};
    if (fc5_ctx_create(device, &ctx) != FC5_OK) {
      std::fprintf(stderr, "[xor] cannot open device %d\n", device);
      return 1;
    }
    if (fc5_table_build(ctx, N, r, nullptr, 0, 1) != FC5_OK || fc5_table_save(ctx, out.c_str()) != FC5_OK ||
        fc5_table_info(ctx, &h) != FC5_OK) {
      std::fprintf(stderr, "[xor] %s\n", fc5_last_error(ctx));
      fc5_ctx_destroy(ctx);
      return 1;
    }
    fc5_ctx_destroy(ctx);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const double gb = packed_bytes(h.m_cells, h.r) / 1e9;
    std::fprintf(stderr, "[xor] N=%d keys=%llu cells=%llu packed=%.2f GB r=%d seed=0x%llx build=%.0f ms (device)\n",
                 N, (unsigned long long)h.n_keys, (unsigned long long)h.m_cells, gb, r,
                 (unsigned long long)h.mix_seed, ms);
    std::printf("[xor_build_save] wrote %s  packed=%.2f GB  cells=%llu  r=%u  N=%u\n", out.c_str(), gb,
                (unsigned long long)h.m_cells, h.r, h.N);
    return 0;
  }
  XorTable t;
  std::string msg;
  if (!build_pairs_cpu(N, (u32)r, {}, t, msg)) { std::fprintf(stderr, "[xor] %s\n", msg.c_str()); return 1; }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "[xor] N=%d keys=%llu cells=%llu packed=%.2f GB r=%d seed=0x%llx build=%.0f ms\n", N,
               (unsigned long long)t.hdr.n_keys, (unsigned long long)t.hdr.m_cells, t.packed.size() / 1e9, r,
               (unsigned long long)t.hdr.mix_seed, ms);
  if (!save_bin(out, t, msg)) { std::fprintf(stderr, "[xor] %s\n", msg.c_str()); return 1; }
  std::printf("[xor_build_save] wrote %s  packed=%.2f GB  cells=%llu  r=%u  N=%u\n", out.c_str(),
              t.packed.size() / 1e9, (unsigned long long)t.hdr.m_cells, t.hdr.r, t.hdr.N);
  return 0;
}
