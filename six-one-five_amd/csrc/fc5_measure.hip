// fc5_measure.hip -- microbenchmarks that set the rooflines of the peel.
//
// k_probe_only   full xor-filter queries of pseudo-random keys against the
//                resident table, every lane a query: the best probe rate the
//                memory system allows at this table size and cell width.
// k_gather_raw   one random aligned 8-byte load per lane over the table's
//                cells: the raw random-gather rate.
// k_copy_u64     a streaming copy of known size, to scale the HBM byte
//                counters (FETCH_SIZE, WRITE_SIZE) of a profiler run.
#include <hip/hip_runtime.h>

#include "fc5_kernels.hpp"

namespace fc5 {

__device__ __forceinline__ u64 m_mix64(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

__global__ __launch_bounds__(256) void k_probe_only(XorDev x, u64 n, u64 salt, unsigned long long* sink) {
  u64 acc = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u64 key = m_mix64(i * 0x9E3779B97F4A7C15ULL + salt);
    const u64 h = m_mix64(key + x.seed);
    const u64 i0 = mod_fast(h, x.md);
    const u64 i1 = mod_fast(m_mix64(h ^ 0x9E3779B97F4A7C15ULL), x.md) + x.block;
    const u64 i2 = mod_fast(m_mix64(h ^ 0xBF58476D1CE4E5B9ULL), x.md) + 2 * x.block;
    acc += ((load_cell_policy(x.cells + i0, x.policy) ^ load_cell_policy(x.cells + i1, x.policy) ^
             load_cell_policy(x.cells + i2, x.policy)) == (h & x.mask))
               ? 1
               : 0;
  }
  if (acc) atomicAdd(sink, (unsigned long long)acc);
}

// md divides by the cell count, so every index is below it
__global__ __launch_bounds__(256) void k_gather_raw(const u64* cells, ModDiv md, u64 n, u64 salt, u32 policy,
                                                    unsigned long long* sink) {
  u64 acc = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    acc ^= load_cell_policy(cells + mod_fast(m_mix64(i + salt), md), policy);
  if (acc == 0x5A5A5A5A5A5A5A5AULL) atomicAdd(sink, 1ULL);
}

__global__ __launch_bounds__(256) void k_copy_u64(const u64* __restrict__ src, u64* __restrict__ dst, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    dst[i] = src[i];
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

}  // namespace fc5
