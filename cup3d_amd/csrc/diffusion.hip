// Implicit diffusion on the device: the block preconditioner and the
// velocity gather/scatter around the three solves of one step
// (AdvectionDiffusionImplicit.h:21-525, AdvectionDiffusionImplicit.cpp:
// 354-419). The operator itself is k_spmv with HELM (krylov.h).
//
// Preconditioner. The reference runs, per block, CG on the local matrix
// A_loc(sigma) = 6-neighbour adjacency + (-6 - h^2 sigma) I with zero
// padding, sigma = 1 / (nu dt), to 1e-7 of its initial residual
// (DiffusionSolverAMRKernels.cpp:149-213), i.e. it approximates
// z_b = h_b^-1 A_loc(sigma_b)^-1 p_b, a matrix that differs per level. That
// matrix is diagonalised by the 3-D discrete sine transform for every sigma:
//   A_loc(sigma)^-1 = S diag(1 / (l_a + l_b + l_c - h^2 sigma)) S,
//   S = S1 (x) S1 (x) S1,  S1[k][i] = sqrt(2/9) sin(pi (k+1)(i+1) / 9),
//   l_k = 2 cos(pi (k+1) / 9) - 2,
// S1 symmetric and orthogonal. So the exact block inverse of any level is
// three separable 8-point transforms, a per-mode scaling and three more
// transforms: 96 flops per cell instead of a 512 x 512 product, no per-level
// matrices, and the kernel is bound by HBM like the vector updates it is
// fused with.
//
// One wavefront per block, lane = (ix, iy), each lane holding its z column:
// the z transform runs in registers, the y and x transforms through this
// wavefront's own LDS slice (rows padded to 9 doubles so the x-transform
// reads of the eight rows fall in distinct banks), so the wavefront
// synchronises only with itself.

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "diffusion.h"
#include "krylov.h"

namespace cup3d {
namespace {

constexpr int DL_ROW = 9;               // padded row (ix) stride, values
constexpr int DL_PLANE = BS * DL_ROW;   // one z plane (or z mode)
constexpr int DL_WAVE = BS * DL_PLANE;  // one wavefront's slice

// out[k] = sum_i S[k][i] in[i] (S symmetric: also the inverse transform);
// the table index is the same in every lane (scalar loads)
template <typename T>
__device__ __forceinline__ void dst8(const T* __restrict__ S, const T (&in)[BS], T (&out)[BS]) {
#pragma unroll
  for (int k = 0; k < BS; ++k) {
    T acc = 0;
#pragma unroll
    for (int i = 0; i < BS; ++i) acc = fma(S[k * BS + i], in[i], acc);
    out[k] = acc;
  }
}

// Transform along y (AXIS 1) or x (AXIS 0) of the eight planes held in the
// slice L: this lane (a, c) = (lane & 7, lane >> 3) produces, for every z
// plane k, the value at output index c (y) or a (x); row holds S[c][*] or
// S[a][*] in registers.
template <int AXIS, typename T>
__device__ __forceinline__ void dst_lds(const T (&row)[BS], const T* L, int lane, T (&out)[BS]) {
  const int a = lane & 7, c = lane >> 3;
#pragma unroll
  for (int k = 0; k < BS; ++k) {
    T acc = 0;
#pragma unroll
    for (int j = 0; j < BS; ++j) {
      const T v = AXIS == 1 ? L[k * DL_PLANE + j * DL_ROW + a] : L[k * DL_PLANE + c * DL_ROW + j];
      acc = fma(row[j], v, acc);
    }
    out[k] = acc;
  }
}

template <typename T>
__device__ __forceinline__ void to_lds(T* L, int lane, const T (&v)[BS]) {
  const int a = lane & 7, c = lane >> 3;
#pragma unroll
  for (int k = 0; k < BS; ++k) L[k * DL_PLANE + c * DL_ROW + a] = v[k];
}

// z = h^-1 A_loc(sigma_b)^-1 v for the column v of this lane (in place).
// tab: S1 (64 values, row-major) then l_0..l_7.
template <typename T>
__device__ __forceinline__ void dst_apply(const T* __restrict__ tab, T* L, int lane, T h,
                                          T sigma, T (&v)[BS]) {
  const int kx = lane & 7, ky = lane >> 3;
  T rx[BS], ry[BS];
#pragma unroll
  for (int j = 0; j < BS; ++j) {
    rx[j] = tab[kx * BS + j];
    ry[j] = tab[ky * BS + j];
  }
  const T base = tab[BS * BS + kx] + tab[BS * BS + ky] - h * h * sigma;
  T w[BS];
  dst8(tab, v, w);  // z transform: w[kz] at (ix, iy)
  to_lds(L, lane, w);
  wave_sync();
  dst_lds<1>(ry, L, lane, v);  // y: v[kz] at (ix, ky)
  wave_sync();
  to_lds(L, lane, v);
  wave_sync();
  dst_lds<0>(rx, L, lane, w);  // x: w[kz] at (kx, ky)
#pragma unroll
  for (int k = 0; k < BS; ++k) w[k] = w[k] / (h * (base + tab[BS * BS + k]));
  wave_sync();
  to_lds(L, lane, w);
  wave_sync();
  dst_lds<0>(rx, L, lane, v);  // inverse x: v[kz] at (ix, ky)
  wave_sync();
  to_lds(L, lane, v);
  wave_sync();
  dst_lds<1>(ry, L, lane, w);  // inverse y: w[kz] at (ix, iy)
  dst8(tab, w, v);             // inverse z: v[iz]
}

// K1 / K3 of the classical iteration (kernels.hip) with this preconditioner,
// in FP64 (the FP64 forms) or FP32 (the inner solve of the mixed forms):
//   PREC_P  p = r + beta (p - omega nu) [p = r after a restart, rhat = r];
//           z1 = M^-1 p
//   PREC_S  s = r - alpha nu (into r); z2 = M^-1 s
//   PREC_PLAIN  z1 = M^-1 p
template <typename T, int MODE>
__global__ __launch_bounds__(WG_THREADS) void k_precond_dst(int nb, const T* __restrict__ tab,
                                                            const T* __restrict__ hb, T sigma,
                                                            KrylovVectors<T> v,
                                                            SolverState* st) {
  __shared__ T Ls[WG_THREADS / 64][DL_WAVE];
  if (MODE != PREC_PLAIN && st->status != ST_RUNNING) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x * (WG_THREADS / 64) + w;
  if (b >= nb) return;  // no workgroup barrier below: only this wavefront's slice
  const T beta = MODE == PREC_P ? (T)st->beta : (T)0;
  const T omega = MODE == PREC_P ? (T)st->omega : (T)0;
  const T alpha = MODE == PREC_S ? (T)st->alpha : (T)0;
  const int pzero = MODE == PREC_P ? st->p_zero : 0;
  const int restart = MODE == PREC_P ? st->restart_pending : 0;
  const size_t off = (size_t)b * BLEN + lane;
  T c[BS], ra[BS], rb[BS];
#pragma unroll
  for (int iz = 0; iz < BS; ++iz) {
    const size_t i = off + iz * 64;
    if (MODE == PREC_P) {
      ra[iz] = v.r[i];
      rb[iz] = v.nu[i];
      c[iz] = v.p[i];
    } else if (MODE == PREC_S) {
      ra[iz] = v.r[i];
      rb[iz] = v.nu[i];
    } else {
      c[iz] = v.p[i];
    }
  }
#pragma unroll
  for (int iz = 0; iz < BS; ++iz) {
    const size_t i = off + iz * 64;
    if (MODE == PREC_P) {
      // BiCGSTAB.cu:518-522
      c[iz] = pzero ? ra[iz] : fma(beta, fma(-omega, rb[iz], c[iz]), ra[iz]);
      v.p[i] = c[iz];
      if (restart) v.rhat[i] = ra[iz];
    } else if (MODE == PREC_S) {
      c[iz] = fma(-alpha, rb[iz], ra[iz]);  // :546-549
      v.r[i] = c[iz];
    }
  }
  dst_apply<T>(tab, Ls[w], lane, hb[b], sigma, c);
  T* __restrict__ out = MODE == PREC_S ? v.z2 : v.z1;
#pragma unroll
  for (int iz = 0; iz < BS; ++iz) out[off + iz * 64] = c[iz];
}

template <typename T>
void launch_precond_dst_t(hipStream_t s, int mode, int nb, const T* tab, const T* h, T sigma,
                          const KrylovVectors<T>& v, SolverState* st) {
  const int nwg = (nb + WG_THREADS / 64 - 1) / (WG_THREADS / 64);
  switch (mode) {
    case PREC_P: k_precond_dst<T, PREC_P><<<nwg, WG_THREADS, 0, s>>>(nb, tab, h, sigma, v, st); break;
    case PREC_S: k_precond_dst<T, PREC_S><<<nwg, WG_THREADS, 0, s>>>(nb, tab, h, sigma, v, st); break;
    default: k_precond_dst<T, PREC_PLAIN><<<nwg, WG_THREADS, 0, s>>>(nb, tab, h, sigma, v, st); break;
  }
}

// Component c of the velocity into the solver's vectors: u = u*_c and
// a = -h^3 sigma u^n_c (the operator's residual a - A u is then the step's
// rhs, see launch_diffusion_gather).
__global__ __launch_bounds__(WG_THREADS) void k_diff_gather(long long n, int comp,
                                                            const double* __restrict__ h3,
                                                            double sigma,
                                                            const double* __restrict__ vel,
                                                            const double* __restrict__ vel_old,
                                                            double* __restrict__ u,
                                                            double* __restrict__ a) {
  const long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x;
  if (i >= n) return;
  u[i] = vel[3 * i + comp];
  a[i] = -h3[i / BLEN] * sigma * vel_old[3 * i + comp];
}

// u*_c += x (AdvectionDiffusionImplicit.cpp:407-418).
__global__ __launch_bounds__(WG_THREADS) void k_diff_scatter(long long n, int comp,
                                                             const double* __restrict__ x,
                                                             double* __restrict__ vel) {
  const long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x;
  if (i >= n) return;
  vel[3 * i + comp] += x[i];
}

}  // namespace

void dst_tables(double* tab, float* tab32) {
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < BS; ++k) {
    tab[BS * BS + k] = 2.0 * std::cos(pi * (k + 1) / 9.0) - 2.0;
    for (int i = 0; i < BS; ++i)
      tab[k * BS + i] = std::sqrt(2.0 / 9.0) * std::sin(pi * (k + 1) * (i + 1) / 9.0);
  }
  for (int i = 0; i < DST_TABLE; ++i) tab32[i] = (float)tab[i];
}

void launch_precond_dst(hipStream_t s, int mode, int nb, const double* tab, const double* h,
                        double sigma, const KrylovVectors<double>& v, SolverState* st) {
  launch_precond_dst_t<double>(s, mode, nb, tab, h, sigma, v, st);
}

void launch_precond_dst32(hipStream_t s, int mode, int nb, const float* tab, const float* h,
                          const KrylovVectors<float>& v, SolverState* st) {
  launch_precond_dst_t<float>(s, mode, nb, tab, h, 0.0f, v, st);
}

void launch_diffusion_gather(hipStream_t s, long long n, int comp, const double* h3, double sigma,
                             const double* vel, const double* vel_old, double* u, double* a) {
  const int nwg = (int)((n + WG_THREADS - 1) / WG_THREADS);
  k_diff_gather<<<nwg, WG_THREADS, 0, s>>>(n, comp, h3, sigma, vel, vel_old, u, a);
}

void launch_diffusion_scatter(hipStream_t s, long long n, int comp, const double* x,
                              double* vel) {
  const int nwg = (int)((n + WG_THREADS - 1) / WG_THREADS);
  k_diff_scatter<<<nwg, WG_THREADS, 0, s>>>(n, comp, x, vel);
}

}  // namespace cup3d
