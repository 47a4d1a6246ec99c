// Device helpers: XCD-aware workgroup order, deterministic reductions and
// the last-workgroup finaliser used by the fused Krylov kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "poisson_internal.h"

namespace cup3d {

// MI355X dispatches workgroups round-robin over its 8 XCDs (workgroup i runs
// on XCD i % 8). Remap so each XCD walks one contiguous range of blocks: the
// same-level neighbours a block reads then sit in the same L2 (the uniform
// 4.2 M-cell SpMV + dot: 14.4 us against 17.1 us in the plain order). Only
// kernels that read neighbours use it: a pure streaming kernel runs 8 %
// faster in the plain order (K5-shaped probe on 30 M cells: 175 us against
// 189 us), which spreads every XCD's requests over all memory channels.
__device__ __forceinline__ int xcd_remap(int i, int G) {
  const int x = i & 7, j = i >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + j;
}

// Orders this wavefront's LDS stores before its later LDS loads of the same
// words (the LDS executes one wavefront's operations in issue order; the fence
// keeps the compiler from moving accesses across this point). Enough where a
// wavefront reads back only what it wrote itself.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum K values over the 256-thread workgroup in a fixed order; the result is
// valid in thread 0 only.
template <int K>
__device__ __forceinline__ void wg_sum(double (&v)[K], double (*red)[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
  }
  __syncthreads();
}

// Producer side of the split reduction: each workgroup stores its K partial
// sums (plain stores; the kernel boundary makes them visible to the
// finalize kernel that follows on the same stream).
template <int K>
__device__ __forceinline__ void grid_store_partial(double (&v)[K], double* __restrict__ part) {
  __shared__ double red[K][4];
  wg_sum<K>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[(size_t)blockIdx.x * K + k] = v[k];
  }
}

// Consumer side (one workgroup): sums the partials of `nparts` producer
// workgroups in workgroup order; the totals are valid in every thread.
template <int K>
__device__ __forceinline__ void grid_sum_partials(const double* __restrict__ part, int nparts,
                                                  double (&tot)[K]) {
  // U partials per thread are loaded before any is added, so the loads of up
  // to U * blockDim.x partials cost one memory round trip, not U (the
  // producers' partials sit in the L2 of every XCD or in HBM). U * K = 60
  // loads stay within the 63 a wavefront can have outstanding. Each thread
  // adds its partials in index order whatever U is, so U does not change
  // the result.
  constexpr int U = 60 / K;
  __shared__ double red[K][4];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int base = 0; base < nparts; base += U * (int)blockDim.x) {
    double t[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * (int)blockDim.x + (int)threadIdx.x;
#pragma unroll
      for (int k = 0; k < K; ++k) t[u][k] = i < nparts ? part[(size_t)i * K + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] += t[u][k];
  }
  wg_sum<K>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][0] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) tot[k] = red[k][0];
}

}  // namespace cup3d
