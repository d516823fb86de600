// fc5_exact.hip -- the exact sorted pair table on the device.
//
// The reference's exact index (QuadSumIndex, quad_sum.hpp:8-45) holds every
// pair sum a^6 + b^6 with 1 <= a <= b <= N sorted by sum, and equal_sum_range
// answers which pairs make one sum.  Here an entry is a single 128-bit word
// sum * 2^s + a (s = bit length of N): the sum orders the entries, a names the
// pair (b is the exact sixth root of sum - a^6), and a = 0 marks an extra value
// (a planted target).  Row a (b = a..N) is written coalesced at its closed-form
// offset, the array is radix-sorted on its significant bits only (rocPRIM
// onesweep over a double buffer), and a lookup is one binary search per target.
#include <rocprim/device/device_radix_sort.hpp>

#include "fc5_device.hpp"

namespace fc5 {

namespace {

// row a starts after rows 1..a-1 of N - a' + 1 entries each
__global__ __launch_bounds__(256) void k_exact_rows(u128* __restrict__ keys, u32 N, u32 s) {
  for (u32 a = blockIdx.x + 1; a <= N; a += gridDim.x) {
    const u64 am1 = a - 1u;
    const u64 off = am1 * ((u64)N + 1) - am1 * a / 2;
    const u128 pa = p6_u128(a);
    for (u32 b = a + threadIdx.x; b <= N; b += blockDim.x) keys[off + (b - a)] = ((pa + p6_u128(b)) << s) | a;
  }
}

// one thread per target: lower bound of T * 2^s, then the run of entries whose
// value is T; pairs with b > hi are left out, an extra value is reported as
// (job, 0, 0) with pad = 1
__global__ __launch_bounds__(256) void k_exact_lookup(const u128* __restrict__ keys, u64 n, u32 s,
                                                      const PairJob* __restrict__ jobs, u32 nj,
                                                      PairHit* __restrict__ out, u32 cap, u32* count) {
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const PairJob J = jobs[j];
  const u128 T = ((u128)J.T_hi << 64) | J.T_lo;
  if (T == 0 || (T >> (128u - s)) != 0) return;  // no entry can hold T
  const u128 want = T << s;
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  const u128 amask = ((u128)1 << s) - 1;
  for (u64 k = lo; k < n; ++k) {
    const u128 key = keys[k];
    if ((key >> s) != T) break;
    const u32 a = (u32)(key & amask);
    u32 b = 0;
    if (a) {
      b = iroot6_u128(T - p6_u128(a));
      if (b > J.hi) continue;
    }
    const u32 slot = atomicAdd(count, 1u);
    if (slot < cap) out[slot] = PairHit{j, a, b, a ? 0u : 1u};
  }
}

}  // namespace

int launch_exact_rows(u128* keys, u32 N, u32 s, hipStream_t st) {
  const u32 blocks = N < 65536u ? N : 65536u;
  hipLaunchKernelGGL(k_exact_rows, dim3(blocks ? blocks : 1), dim3(256), 0, st, keys, N, s);
  return (int)hipGetLastError();
}

int exact_sort(u128* keys, u128* alt, u64 n, u32 end_bit, void* temp, size_t& temp_bytes, hipStream_t st,
               bool& sorted_in_alt) {
  rocprim::double_buffer<rocprim::uint128_t> db(keys, alt);
  const hipError_t e = rocprim::radix_sort_keys(temp, temp_bytes, db, (size_t)n, 0u, end_bit, st);
  if (temp) sorted_in_alt = db.current() == alt;
  return (int)e;
}

int launch_exact_lookup(const u128* keys, u64 n, u32 s, const PairJob* jobs, u32 nj, PairHit* out, u32 cap,
                        u32* count, hipStream_t st) {
  const u32 blocks = nj ? (nj + 255) / 256 : 1;
  hipLaunchKernelGGL(k_exact_lookup, dim3(blocks), dim3(256), 0, st, keys, n, s, jobs, nj, out, cap, count);
  return (int)hipGetLastError();
}

}  // namespace fc5
