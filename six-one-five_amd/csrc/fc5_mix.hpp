// fc5_mix.hpp -- the xor filter's hash: one definition for host and device.
//
// The pair-table format fixes these (spike/ribbon/mix64.hpp:13-25,
// spike/ribbon/xor_filter.hpp:53-64): the murmur3-style 64-bit finalizer, the
// seeded hash h = mix64(key + seed), and the three cells of a key, one in each
// third of the table: h % block, mix64(h ^ kCellSalt1) % block + block,
// mix64(h ^ kCellSalt2) % block + 2 block.  Every builder, query and
// microbenchmark (host and device) calls these, so the table bytes and the
// false-positive lists cannot drift apart between them.
#pragma once

#include <cstdint>

#include "fc5_moddiv.hpp"

namespace fc5 {

constexpr std::uint64_t kCellSalt1 = 0x9E3779B97F4A7C15ULL;
constexpr std::uint64_t kCellSalt2 = 0xBF58476D1CE4E5B9ULL;

FC5_HD std::uint64_t mix64(std::uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// the three cell indices of a hashed key h (block = m_cells / 3, md divides by block)
FC5_HD void xor_cells(std::uint64_t h, std::uint64_t block, const ModDiv& md, std::uint64_t c[3]) {
  c[0] = mod_fast(h, md);
  c[1] = mod_fast(mix64(h ^ kCellSalt1), md) + block;
  c[2] = mod_fast(mix64(h ^ kCellSalt2), md) + 2 * block;
}

}  // namespace fc5
