// The caller's post-solve mean removal (PressureProjection.cpp:404-431) for
// callers that cannot allocate device memory: HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc. It has no counterpart in
// the reference. Unlike cup3d_poisson_remove_mean it needs no solver handle
// (whose creation allocates): the partial sums of its two-pass reduction
// live in `scratch`, a static device array of this library, sized for
// NOALLOC_MAX_BLOCKS blocks; each GPU holds its own copy until the process
// ends.
//
// The arithmetic is that of cup3d_poisson_remove_mean, summation order
// included: per workgroup of four blocks sum_cells h3 x (8 cells per thread,
// then the workgroup reduction), the partials added in workgroup order by
// one workgroup, m = wsum * (1 / volume), x -= m. The volume sum_b 512 h_b^3
// is summed on the device here and on the host there; its terms are dyadic
// (h_b = h0 / 2^level), so both sums are exact and equal while the levels
// span fewer than 53 binary orders of magnitude of h^3.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>

#include "cup3d_poisson.h"
#include "device_common.h"
#include "krylov.h"

namespace cup3d {
namespace {

constexpr int64_t NOALLOC_PARTS = CUP3D_NOALLOC_MAX_BLOCKS / 4;  // workgroups of 4 blocks
constexpr int64_t NOALLOC_TOTALS = 2 * NOALLOC_PARTS;           // offset of the two totals

// [2 * NOALLOC_PARTS] partial sums (sum h3 x, volume), then the two totals:
// 4 MiB + 16 bytes per GPU.
static __device__ __align__(256) double scratch[2 * NOALLOC_PARTS + 2];

// The device address of `scratch` on each device (the symbol's name is only
// a host-side stand-in), looked up once per device with it current.
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

__global__ __launch_bounds__(WG_THREADS) void k_wsum_vol(long long n,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ h3,
                                                         double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  const long long i0 = (long long)blockIdx.x * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    double xv[8];
    load8(x + i0, xv);
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) s += xv[e];
    const double hb = h3[i0 / BLEN];
    v[0] = hb * s;
    v[1] = 8.0 * hb;  // this thread's 8 cells of volume h^3
  }
  grid_store_partial<2>(v, part);
}

__global__ __launch_bounds__(WG_THREADS) void k_wsum_vol_fin(const double* __restrict__ part,
                                                             int nparts,
                                                             double* __restrict__ tot_out) {
  double tot[2];
  grid_sum_partials<2>(part, nparts, tot);
  if (threadIdx.x == 0) {
    tot_out[0] = tot[0];
    tot_out[1] = tot[1];
  }
}

__global__ __launch_bounds__(WG_THREADS) void k_sub_mean_tot(long long n, double* __restrict__ x,
                                                             const double* __restrict__ tot) {
  const double m = tot[0] * (1.0 / tot[1]);
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    x[i] -= m;
}

int fail(const char* m) {
  set_global_error(m);
  return -1;
}

}  // namespace
}  // namespace cup3d

using namespace cup3d;

extern "C" int cup3d_remove_mean_noalloc(int64_t nblocks, const double* d_h3, double* d_x,
                                         void* stream, double* mean) {
  if (nblocks <= 0) return fail("nblocks must be positive");
  if (nblocks > CUP3D_NOALLOC_MAX_BLOCKS)
    return fail("too many blocks for the static scratch array (CUP3D_NOALLOC_MAX_BLOCKS)");
  if (!d_h3 || !d_x) return fail("null vector");
  if ((uintptr_t)d_x & 15) return fail("device vectors must be 16-byte aligned");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("hipGetDevice failed");
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device ordinal out of range");
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return fail("hipGetSymbolAddress(scratch) failed");
      g_scratch[dev] = static_cast<double*>(p);
    }
    ptr = g_scratch[dev];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = (long long)nblocks * BLEN;
  const int nwg = (int)((n + 4 * BLEN - 1) / (4 * BLEN));
  // every partial and both totals are written before they are read: nothing
  // left from an earlier call is used, and nothing needs clearing
  k_wsum_vol<<<nwg, WG_THREADS, 0, s>>>(n, d_x, d_h3, ptr);
  k_wsum_vol_fin<<<1, WG_THREADS, 0, s>>>(ptr, nwg, ptr + NOALLOC_TOTALS);
  const int nsub = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 8192);
  k_sub_mean_tot<<<nsub, WG_THREADS, 0, s>>>(n, d_x, ptr + NOALLOC_TOTALS);
  if (hipGetLastError() != hipSuccess) return fail("kernel launch failed");
  if (mean) {
    double tot[2];
    if (hipMemcpyAsync(tot, ptr + NOALLOC_TOTALS, sizeof(tot), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail("reading the mean failed");
    *mean = tot[0] * (1.0 / tot[1]);
  }
  return 0;
}
