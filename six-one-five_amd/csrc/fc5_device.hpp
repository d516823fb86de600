// fc5_device.hpp -- device helpers shared by the search kernels (gfx950).
//
// xor-filter query on expanded cells, wavefront reductions
// and lane reads, the LDS residue gate test and an exact sixth root.
#pragma once

#include "fc5_kernels.hpp"
#include "fc5_mix.hpp"

namespace fc5 {

typedef unsigned __int128 u128;

constexpr u32 kGateMod = 124488;  // 504 * 247
constexpr u32 kGateWords = (kGateMod + 31) / 32;

// xor filter query on expanded cells (xor_filter.hpp:53-64,472-483)
__device__ __forceinline__ bool d_query(const XorDev& x, u64 key) {
  if (!x.present) return false;
  const u64 h = mix64(key + x.seed);
  u64 c[3];
  xor_cells(h, x.block, x.md, c);
  const u64 got = load_cell_policy(x.cells + c[0], x.policy) ^ load_cell_policy(x.cells + c[1], x.policy) ^
                  load_cell_policy(x.cells + c[2], x.policy);
  return got == (h & x.mask);
}

__device__ __forceinline__ u64 wave_sum(u64 v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ u32 wave_sum_u32(u32 v) {
  for (int o = 32; o > 0; o >>= 1) v += (u32)__shfl_xor((int)v, o, 64);
  return v;
}

__device__ __forceinline__ u32 wave_max_u32(u32 v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (u32)__shfl_xor((int)v, o, 64));
  return v;
}

__device__ __forceinline__ u32 rdl(u32 v, int lane) {
  return (u32)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ u32 lanes_below(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// (x + n) mod 124488 is in the LDS bitmap; x, n < 124488
__device__ __forceinline__ bool gate_pass(const u32* s_gate, u32 xg, u32 n) {
  u32 t = xg + n;
  t = min(t, t - kGateMod);
  return (s_gate[t >> 5] >> (t & 31u)) & 1u;
}

// (x y) mod 124488 in 32-bit steps for x y < 2^51: the high word of the
// product (below 2^19) times 2^32 mod 124488 = 6808, plus the reduced low
// word, stays below 2^32
__device__ __forceinline__ u32 mulmod_gate(u32 x, u32 y) {
  const u64 p = (u64)x * y;
  const u32 r = (u32)p % kGateMod + (u32)(p >> 32) * 6808u;
  return r % kGateMod;
}

// 124488 - (c^6 mod 124488), in [1, 124488]: the n of gate_pass for the
// residual T - c^6, computed in registers (c < 2^25)
__device__ __forceinline__ u32 neg_p6_gate(u32 c) {
  const u32 c2 = mulmod_gate(c, c), c3 = mulmod_gate(c2, c);
  return kGateMod - mulmod_gate(c3, c3);
}

// X mod 124488 for a 128-bit X
__device__ __forceinline__ u32 mod_gate(u128 X) {
  const u64 xh = (u64)(X >> 64), xl = (u64)X;
  return (u32)(((xh % kGateMod) * (u64)(18446744073709551615ULL % kGateMod + 1) + xl % kGateMod) % kGateMod);
}

__device__ __forceinline__ u128 p6_u128(u32 v) {
  const u64 v3 = (u64)v * v * v;
  return (u128)v3 * v3;
}

// largest a with a^6 <= X, for X < 2^127: the root is then below 2.36e6, so
// (a+1)^3 stays below 2^64 in the fix-up loops
__device__ __forceinline__ u32 iroot6_u128(u128 X) {
  const double x = (double)(u64)(X >> 64) * 18446744073709551616.0 + (double)(u64)X;
  u32 a = (u32)cbrt(sqrt(x));
  while (p6_u128(a + 1) <= X) ++a;
  while (a > 0 && p6_u128(a) > X) --a;
  return a;
}

// smallest c >= 1 with k * c^6 >= X (k = 3 or 4, X > 0)
__device__ __forceinline__ u32 min_core(u128 X, u32 k) {
  const double x = (double)(u64)(X >> 64) * 18446744073709551616.0 + (double)(u64)X;
  u32 l = (u32)cbrt(sqrt(x / (double)k));
  if (l < 1) l = 1;
  while (l > 1 && p6_u128(l - 1) * k >= X) --l;
  while (p6_u128(l) * k < X) ++l;
  return l;
}

}  // namespace fc5
