// fc5_verify.hip -- exact check of filter survivors on the device.
//
// k_pair_search: for each job (an exact 128-bit target T and a bound hi) find
// every pair a <= b <= hi with a^6 + b^6 = T.  One wavefront per job; the
// lanes run over b in [least b with 2 b^6 >= T, min(iroot6(T), hi)] (b is the
// larger term) and test whether T - b^6 is a sixth power with an exact 128-bit
// root.  A survivor of the xor filter is a real pair sum exactly when this
// finds a pair; the pairs are what the reference recovers on the host
// (PairRecover + exact compare, fourcore_xor_store.hpp:23-51,
// fourcore_cls5_gpu_v4.cu:517-522).
#include "fc5_device.hpp"

namespace fc5 {

__global__ __launch_bounds__(256) void k_pair_search(const PairJob* __restrict__ jobs, u32 n,
                                                     PairHit* __restrict__ out, u32 cap, u32* count) {
  const u32 lane = threadIdx.x & 63u;
  const u32 waves = gridDim.x * (blockDim.x >> 6);
  for (u32 j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < n; j += waves) {
    const PairJob J = jobs[j];
    const u128 T = ((u128)J.T_hi << 64) | J.T_lo;
    if (T == 0 || J.hi == 0) continue;
    const u32 top = min(iroot6_u128(T), J.hi);
    const u32 lo = min_core(T, 2u);  // b >= a  <=>  2 b^6 >= T
    for (u32 bb = lo; bb <= top; bb += 64u) {
      const u32 b = bb + lane;
      if (b <= top) {
        const u128 r = T - p6_u128(b);
        const u32 a = iroot6_u128(r);
        if (a >= 1 && a <= b && p6_u128(a) == r) {
          const u32 s = atomicAdd(count, 1u);
          if (s < cap) out[s] = PairHit{j, a, b, 0u};
        }
      }
      if (top - bb < 64u) break;
    }
  }
}

int launch_pair_search(const PairJob* jobs, u32 n, PairHit* out, u32 cap, u32* count, hipStream_t s) {
  const u32 blocks = n ? (n + 3) / 4 : 1;
  hipLaunchKernelGGL(k_pair_search, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, s, jobs, n, out, cap, count);
  return (int)hipGetLastError();
}

}  // namespace fc5
