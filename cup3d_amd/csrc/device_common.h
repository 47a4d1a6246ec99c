// Device helpers: XCD-aware workgroup order, deterministic reductions and
// the last-workgroup finaliser used by the fused Krylov kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "poisson_internal.h"

namespace cup3d {

// MI355X dispatches workgroups round-robin over its 8 XCDs (workgroup i runs
// on XCD i % 8). Remap so each XCD walks one contiguous range of blocks: the
// same-level neighbours a block reads then sit in the same L2.
__device__ __forceinline__ int xcd_remap(int i, int G) {
  const int x = i & 7, j = i >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + j;
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

// Two-pass deterministic reduction across the grid. Every workgroup publishes
// its partial sums; the workgroup that arrives last adds all partials in
// workgroup-id order (independent of arrival order) and returns true in all
// its threads with the totals in `tot`. The ticket counter is reset by the
// last workgroup, so the next kernel finds it at zero.
//
// Cross-XCD visibility without cache flushes: partials are written and read
// with agent-scope atomic stores/loads (coherent at the device level, they do
// not go through a non-coherent L2 copy), and the store is complete (s_waitcnt)
// before the ticket increment is issued. A device-scope release fence per
// workgroup would instead write back the whole L2 of the XCD every time.
template <int K>
__device__ bool grid_reduce_last(double (&v)[K], double* __restrict__ part,
                                 unsigned int* counter, double (&tot)[K]) {
  __shared__ double red[K][4];
  __shared__ int am_last;
  const int nwg = gridDim.x;
  wg_sum<K>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      __hip_atomic_store(&part[(size_t)blockIdx.x * K + k], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    const unsigned t =
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    am_last = (t == (unsigned)(nwg - 1));
  }
  __syncthreads();
  if (!am_last) return false;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < nwg; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      acc[k] += __hip_atomic_load(&part[(size_t)i * K + k], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  }
  wg_sum<K>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k][0] = acc[k];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) tot[k] = red[k][0];
  return true;
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
  __shared__ double red[K][4];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += part[(size_t)i * K + k];
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
