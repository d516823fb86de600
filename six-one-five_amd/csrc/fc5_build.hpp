// fc5_build.hpp -- device construction of the xor pair table (see fc5_build.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace fc5 {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

struct StoreHeaderPod {  // StoreHeader v2 layout
  u32 magic, version, kind, r, w, N;
  u64 n_keys, m_cells, mix_seed;
  u32 shard_count, shard_index, shard_mode, reserved;
};
static_assert(sizeof(StoreHeaderPod) == 64, "header is 64 bytes");

struct GpuBuilt {
  StoreHeaderPod hdr{};
  u64* cells = nullptr;  // expanded cells (device), owned by the caller
  u64 rounds = 0;        // peel rounds
};

// pair keys for 1<=a<=b<=N plus extra keys; returns 0 or -1 with err set
int gpu_build_pairs(int N, u32 r, const u64* extra, size_t n_extra, hipStream_t s, GpuBuilt& out,
                    std::string& err);
// r-bit packing of m expanded cells into out_dev (ceil(m/8)*r bytes writable)
int gpu_pack_cells(const u64* cells, u64 m, u32 r, u8* out_dev, hipStream_t s);

}  // namespace fc5
