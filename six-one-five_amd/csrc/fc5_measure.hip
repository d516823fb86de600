// fc5_measure.hip -- microbenchmarks that set the rooflines of the peel.
//
// k_probe_only   full xor-filter queries of pseudo-random keys against the
//                resident table, every lane a query: the best probe rate the
//                memory system allows at this table size and cell width.
// k_gather_raw   one random aligned 8-byte load per lane over the first m
//                cells of the table: the raw random-gather rate, and, with m
//                swept from a few MB to the whole table, how much the L2 and
//                the 256 MB memory-side cache (MALL) lift it.
// k_copy_u64     a streaming copy of known size, to scale the HBM byte
//                counters (FETCH_SIZE, WRITE_SIZE) of a profiler run.
// k_read_stream  a read-only streaming pass (16-byte non-temporal loads, 8 in
//                flight per lane): the sequential read ceiling of HBM, and a
//                second calibration of FETCH_SIZE with no writes beside it.
// k_valu<K>      integer VALU issue rate: chains of v_mad_u64_u32 (K = 1),
//                128-bit add-with-carry chains (K = 2), 32-bit operations
//                (K = 0) and the class-5 kernel's mix (K = 3: one 64-bit
//                multiply-add per 13 32-bit operations, 7% INT64).
#include <hip/hip_runtime.h>

#include "fc5_kernels.hpp"
#include "fc5_mix.hpp"

namespace fc5 {

__global__ __launch_bounds__(256) void k_probe_only(XorDev x, u64 n, u64 salt, unsigned long long* sink) {
  u64 acc = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u64 key = mix64(i * 0x9E3779B97F4A7C15ULL + salt);
    const u64 h = mix64(key + x.seed);
    u64 c[3];
    xor_cells(h, x.block, x.md, c);
    acc += ((load_cell_policy(x.cells + c[0], x.policy) ^ load_cell_policy(x.cells + c[1], x.policy) ^
             load_cell_policy(x.cells + c[2], x.policy)) == (h & x.mask))
               ? 1
               : 0;
  }
  if (acc) atomicAdd(sink, (unsigned long long)acc);
}

// md divides by the cell count m, so every index is below it
__global__ __launch_bounds__(256) void k_gather_raw(const u64* cells, ModDiv md, u64 n, u64 salt, u32 policy,
                                                    unsigned long long* sink) {
  u64 acc = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    acc ^= load_cell_policy(cells + mod_fast(mix64(i + salt), md), policy);
  if (acc == 0x5A5A5A5A5A5A5A5AULL) atomicAdd(sink, 1ULL);
}

__global__ __launch_bounds__(256) void k_copy_u64(const u64* __restrict__ src, u64* __restrict__ dst, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// n16 16-byte words; each lane reads words i + k S, k < 8, per step (S =
// threads of the grid), so a wave's eight non-temporal loads are eight
// coalesced 1 KiB runs (measured the fastest of the strided and per-block
// chunked forms with 4 or 8 loads in flight, plain or non-temporal)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_read_stream(const u32x4* __restrict__ src, u64 n16,
                                                     unsigned long long* sink) {
  constexpr int U = 8;
  const u64 S = (u64)gridDim.x * blockDim.x;
  u32 acc = 0;
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * S < n16; i += U * S) {
    u32x4 v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = __builtin_nontemporal_load(src + i + k * S);
#pragma unroll
    for (int k = 0; k < U; ++k) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
  }
  for (; i < n16; i += S) {
    const u32x4 a = src[i];
    acc ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (acc == 0x5A5A5A5Au) atomicAdd(sink, 1ULL);
}

// Eight independent dependency chains per lane, `iters` steps each.  The step
// of each kind (one chain):
//   K = 0: x = (x ^ y) + z (32-bit; one v_xad_u32 or v_xor + v_add)
//   K = 1: X = lo32(X) * m + Y (one v_mad_u64_u32)
//   K = 2: A = A + B, B = B + A over 128 bits (v_add_co_u32 + 3 v_addc_co_u32 each)
//   K = 3: X = lo32(X) * m + Y, then 13 32-bit steps (x ^ y) + z on its words
// The chains depend on the lane index and the loop never folds (each step
// consumes the previous one's result), so every step issues.
template <int K>
__global__ __launch_bounds__(256) void k_valu(u32 iters, u32 seed, unsigned long long* sink) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  u64 X[8], Y[8];
  u32 y[8], z[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    X[j] = ((u64)(t * 0x9E3779B9u + j) << 32) | (seed ^ (u32)j * 0x85EBCA6Bu);
    Y[j] = X[j] * 0xC2B2AE3D27D4EB4FULL + (u64)j;
    y[j] = (u32)(Y[j] >> 17) | 1u;
    z[j] = (u32)(Y[j] >> 40) * 3u + 7u;
  }
  const u32 m = seed | 0x10001u;
  for (u32 it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (K == 0) {
        u32 x = (u32)X[j];
        x = (x ^ y[j]) + z[j];
        X[j] = (X[j] & 0xFFFFFFFF00000000ULL) | x;
      } else if (K == 1) {
        X[j] = (u64)(u32)X[j] * m + Y[j];
      } else if (K == 2) {
        if (j & 1) continue;  // chains j and j + 1 hold one pair (A, B): A += B, then B += A
        u128 A = ((u128)X[j] << 64) | Y[j], Bv = ((u128)X[j + 1] << 64) | Y[j + 1];
        A += Bv;
        Bv += A;
        X[j] = (u64)(A >> 64);
        Y[j] = (u64)A;
        X[j + 1] = (u64)(Bv >> 64);
        Y[j + 1] = (u64)Bv;
      } else {
        X[j] = (u64)(u32)X[j] * m + Y[j];
        u32 lo = (u32)X[j], hi = (u32)(X[j] >> 32);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
          lo = (lo ^ y[j]) + hi;
          hi = (hi ^ z[j]) + lo;
        }
        lo = (lo ^ y[j]) + z[j];
        X[j] = ((u64)hi << 32) | lo;
      }
    }
  }
  u64 acc = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc ^= X[j] ^ (K == 2 ? Y[j] : 0);
  if (acc == 0x5A5A5A5A5A5A5A5AULL) atomicAdd(sink, 1ULL);
}

int launch_probe_only(const XorDev& x, u64 n, u64 salt, unsigned long long* sink, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_only, dim3(grid), dim3(256), 0, s, x, n, salt, sink);
  return (int)hipGetLastError();
}

int launch_gather_raw(const u64* cells, u64 m, u64 n, u64 salt, u32 policy, unsigned long long* sink, int grid,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_gather_raw, dim3(grid), dim3(256), 0, s, cells, make_moddiv(m), n, salt, policy, sink);
  return (int)hipGetLastError();
}

int launch_copy_u64(const u64* src, u64* dst, u64 n, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_u64, dim3(grid), dim3(256), 0, s, src, dst, n);
  return (int)hipGetLastError();
}

int launch_read_stream(const void* src, u64 bytes, unsigned long long* sink, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_read_stream, dim3(grid), dim3(256), 0, s, (const u32x4*)src, bytes / 16, sink);
  return (int)hipGetLastError();
}

int launch_valu(int kind, u32 iters, u32 seed, unsigned long long* sink, int grid, hipStream_t s) {
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_valu<0>, dim3(grid), dim3(256), 0, s, iters, seed, sink); break;
    case 1: hipLaunchKernelGGL(k_valu<1>, dim3(grid), dim3(256), 0, s, iters, seed, sink); break;
    case 2: hipLaunchKernelGGL(k_valu<2>, dim3(grid), dim3(256), 0, s, iters, seed, sink); break;
    case 3: hipLaunchKernelGGL(k_valu<3>, dim3(grid), dim3(256), 0, s, iters, seed, sink); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

}  // namespace fc5
