// Mixed-precision form of the solve: an FP64 iterative-refinement loop around
// an FP32 classical BiCGSTAB whose block-Jacobi preconditioner runs on the
// FP16 (or BF16) matrix cores with FP32 accumulation.
//
//   outer (FP64, on the device):  r = b - A x        (true residual, FP64 dot)
//                                 stop if |r| <= tol
//                                 solve A d = r / |r| in FP32 to inner_tol
//                                 x += |r| d
//   inner: the five fused kernels of the FP64 solver instantiated for float,
//          with K1/K3 replaced by the half-precision GEMM below.
//
// The half-precision GEMM. P_inv entries span [3.4e-6, 0.237] in magnitude;
// scaled by 2^14 they sit in FP16's normal range (max 3.9e3, min 5.6e-2).
// Each block's input vector is scaled by a power of two so its largest
// entry lies in [0.5, 1) before the cast, and both scales are undone in the
// FP32 epilogue together with h_b^-1. BF16 needs no P scaling (FP32 exponent
// range) but uses the same per-block input scaling.
// MFMA: v_mfma_f32_16x16x32_{f16,bf16}; A = 16 blocks x 32 cells of the
// input (LDS), B = 32 x 16 of P_inv (fragment-packed in global memory, 1 KiB
// per fragment), D = 16 blocks x 16 output cells (FP32).
//
// Split input (default). A single FP16 pass rounds every input vector to
// 11 bits in a data-dependent way, so M^-1 is no longer a fixed linear
// operator (BiCGSTAB then loses its bi-orthogonality and stagnates over
// hundreds of iterations). P_inv itself is rounded once (a fixed, linear
// perturbation of the preconditioner, harmless), but each block's input is
// carried as hi + lo with lo = v - hi (the block scaling keeps |v| <= 1, so
// lo's absolute error is ~2^-24 of the block maximum even where it is
// subnormal) and the GEMM runs two MFMA passes, P*hi + P*lo, into one FP32
// accumulator: M^-1 is then linear to FP32 precision while every product
// still runs on the FP16/BF16 matrix cores.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "device_common.h"
#include "kernels.h"
#include "krylov.h"
#include "diffusion.h"
#include "mixed.h"

namespace cup3d {

// float4 access to the vectors the preconditioner kernels stream through
// once: nontemporal (NT) for vectors too large to stay in the Infinity Cache
// between kernels (STREAM_MIN_BYTES).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ float4 ld4(const float* base, size_t i4) {
  if constexpr (NT) {
    const nt_f4 q = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(base) + i4);
    return make_float4(q.x, q.y, q.z, q.w);
  } else {
    return reinterpret_cast<const float4*>(base)[i4];
  }
}
template <bool NT>
__device__ __forceinline__ void st4(float* base, size_t i4, const float4& v) {
  if constexpr (NT) {
    const nt_f4 q = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(q, reinterpret_cast<nt_f4*>(base) + i4);
  } else {
    reinterpret_cast<float4*>(base)[i4] = v;
  }
}


typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf168_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename H>
struct HalfVec;
template <>
struct HalfVec<_Float16> {
  typedef half8_t type;
};
template <>
struct HalfVec<__bf16> {
  typedef bf168_t type;
};

__device__ __forceinline__ float4_t mfma16(half8_t a, half8_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4_t mfma16(bf168_t a, bf168_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Accumulator layout of v_mfma_f32_16x16x32_*: register q of lane l holds
// D[4*(l/16) + q][l%16].
__device__ __forceinline__ int acc_row(int g, int q) { return 4 * g + q; }

constexpr int HLD = BLEN + 8;  // LDS row stride in halfs (1040 bytes)
constexpr float P16_SCALE = 16384.0f;
constexpr int GW = 8;          // waves per GEMM workgroup
constexpr int GT = GW * 64;    // threads per GEMM workgroup

// One workgroup applies M^-1 to NBW consecutive blocks.
//  phase 1: wave w owns blocks [w NBW/8, (w+1) NBW/8); every load of the
//           fused vector update is issued before the first one is used
//           (float4, contiguous), then the per-block power-of-two scale and
//           the hi/lo cast into LDS.
//  phase 2: wave w owns output cells [64 w, 64 w + 64) of all NBW blocks:
//           16 k-steps of 32 cells, each 4 (cell tiles) x NBW/16 (block
//           tiles) x (1 or 2 passes) MFMAs; the P_inv fragments of the next
//           k-step are prefetched from L2 while the current ones are used.
//  phase 3: FP32 epilogue, scaled by h_b^-1 and the two scales.

// Scalars of the fused vector update, read once per workgroup.
struct PrecScalars {
  float beta, omega, alpha;
  int pzero, restart;
};

template <int MODE>
__device__ __forceinline__ PrecScalars prec_scalars(const SolverState* st) {
  PrecScalars c;
  c.beta = MODE == PREC_P ? (float)st->beta : 0.f;
  c.omega = MODE == PREC_P ? (float)st->omega : 0.f;
  c.alpha = MODE == PREC_S ? (float)st->alpha : 0.f;
  c.pzero = MODE == PREC_P ? st->p_zero : 0;
  c.restart = MODE == PREC_P ? st->restart_pending : 0;
  return c;
}

// Phase 1 for one wave of NWP filling waves: blocks [w NBW/NWP, (w+1) NBW/NWP)
// of the group starting at block j0. Element u of a lane: block jw + u/2,
// cells 256 (u%2) + 4 l .. +3. precond_load issues every load of the fused
// vector update; precond_convert (after them) stores the updated vector,
// finds the per-block power-of-two scale and casts hi/lo into LDS
// (A0 = hi, A1 = lo).
template <int NBW, int NWP>
struct FillRegs {
  static constexpr int NPB = NBW / NWP;  // blocks per wave
  static constexpr int F4 = 2 * NPB;     // float4 per lane per vector (a block is 2 x 64 float4)
  static_assert(NBW % NWP == 0, "block group must fill the waves");
  float4 ra[F4], rb[F4], rc[F4];
};

template <int MODE, int NBW, int NWP, bool NT>
__device__ __forceinline__ void precond_load(int nb, int j0, int w, int l,
                                             const KrylovVectors<float>& v,
                                             const float* __restrict__ pin,
                                             FillRegs<NBW, NWP>& R) {
  constexpr int F4 = FillRegs<NBW, NWP>::F4;
  const int jw = w * FillRegs<NBW, NWP>::NPB;
#pragma unroll
  for (int u = 0; u < F4; ++u) {
    const int b = j0 + jw + u / 2;
    const size_t i4 = ((size_t)b * BLEN + 256 * (u & 1) + 4 * l) / 4;
    if (b < nb) {
      if (MODE == PREC_P) {
        R.ra[u] = ld4<NT>(v.r, i4);
        R.rb[u] = ld4<NT>(v.nu, i4);
        R.rc[u] = ld4<NT>(v.p, i4);
      } else if (MODE == PREC_S) {
        R.ra[u] = ld4<NT>(v.r, i4);
        R.rb[u] = ld4<NT>(v.nu, i4);
      } else {
        R.ra[u] = ld4<NT>(pin, i4);
      }
    } else {
      R.ra[u] = R.rb[u] = R.rc[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

template <int MODE, typename H, int NBW, bool SPLIT, int NWP, bool NT>
__device__ __forceinline__ void precond_convert(int nb, int j0, int w, int l, float p_unscale,
                                                const float* __restrict__ invh,
                                                const KrylovVectors<float>& v,
                                                const PrecScalars& c, FillRegs<NBW, NWP>& R,
                                                H* A0, H* A1, float* sc) {
  typedef H H4 __attribute__((ext_vector_type(4)));
  constexpr int NPB = FillRegs<NBW, NWP>::NPB, F4 = FillRegs<NBW, NWP>::F4;
  const int jw = w * NPB;
  float amax[NPB];
#pragma unroll
  for (int jj = 0; jj < NPB; ++jj) amax[jj] = 0.f;
#pragma unroll
  for (int u = 0; u < F4; ++u) {
    const int b = j0 + jw + u / 2;
    const size_t i4 = ((size_t)b * BLEN + 256 * (u & 1) + 4 * l) / 4;
    float4 val = R.ra[u];
    if (MODE == PREC_P) {
      if (!c.pzero) {
        val.x = fmaf(c.beta, fmaf(-c.omega, R.rb[u].x, R.rc[u].x), R.ra[u].x);
        val.y = fmaf(c.beta, fmaf(-c.omega, R.rb[u].y, R.rc[u].y), R.ra[u].y);
        val.z = fmaf(c.beta, fmaf(-c.omega, R.rb[u].z, R.rc[u].z), R.ra[u].z);
        val.w = fmaf(c.beta, fmaf(-c.omega, R.rb[u].w, R.rc[u].w), R.ra[u].w);
      }
      if (b < nb) {
        st4<NT>(v.p, i4, val);
        if (c.restart) st4<NT>(v.rhat, i4, R.ra[u]);
      }
    } else if (MODE == PREC_S) {
      val.x = fmaf(-c.alpha, R.rb[u].x, R.ra[u].x);
      val.y = fmaf(-c.alpha, R.rb[u].y, R.ra[u].y);
      val.z = fmaf(-c.alpha, R.rb[u].z, R.ra[u].z);
      val.w = fmaf(-c.alpha, R.rb[u].w, R.ra[u].w);
      if (b < nb) st4<NT>(v.r, i4, val);
    }
    R.ra[u] = val;
    amax[u / 2] = fmaxf(amax[u / 2], fmaxf(fmaxf(fabsf(val.x), fabsf(val.y)),
                                           fmaxf(fabsf(val.z), fabsf(val.w))));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int jj = 0; jj < NPB; ++jj) amax[jj] = fmaxf(amax[jj], __shfl_xor(amax[jj], o, 64));
#pragma unroll
  for (int u = 0; u < F4; ++u) {
    const int jj = u / 2;
    int ex = 0;
    if (amax[jj] > 0.f) frexpf(amax[jj], &ex);
    const float s = amax[jj] > 0.f ? ldexpf(1.0f, -ex) : 1.0f;
    const float x[4] = {R.ra[u].x * s, R.ra[u].y * s, R.ra[u].z * s, R.ra[u].w * s};
    H4 hi, lo;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      hi[k] = (H)x[k];
      lo[k] = (H)(x[k] - (float)hi[k]);
    }
    const int o = (jw + jj) * HLD + 256 * (u & 1) + 4 * l;
    *reinterpret_cast<H4*>(&A0[o]) = hi;
    if (SPLIT) *reinterpret_cast<H4*>(&A1[o]) = lo;
    if ((u & 1) == 0 && l == 0) {
      const int b = j0 + jw + jj;
      sc[jw + jj] = b < nb ? invh[b] * p_unscale / s : 0.f;
    }
  }
}

// Phases 2 and 3 for GEMM wave w (of GW): output cells [64 w, 64 w + 64) of
// all NBW blocks of the group staged in A0/A1; 16 k-steps of 32 cells, each
// 4 (cell tiles) x NBW/16 (block tiles) x (1 or 2 passes) MFMAs, the P_inv
// fragments of the next k-step prefetched from L2 while the current ones are
// used; FP32 epilogue scaled by h_b^-1 and the two scales.
template <typename H, int NBW, bool SPLIT, bool NT>
__device__ __forceinline__ void precond_gemm(int nb, int j0, int w, int l,
                                             const H* __restrict__ Ppk, const H* A0,
                                             const H* A1, const float* sc,
                                             float* __restrict__ out) {
  typedef typename HalfVec<H>::type HV;
  constexpr int JT = NBW / 16;
  static_assert(NBW % 16 == 0, "block group must fill the tiles");
  const int lr = l & 15, g = l >> 4;
  float4_t acc[4][JT];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[t][jt] = (float4_t){0.f, 0.f, 0.f, 0.f};
  // fragment (it, kt) of the packed P_inv: 64 lanes x 8 values at ((it 16 + kt) 64 + l) 8
  const HV* __restrict__ pb = reinterpret_cast<const HV*>(Ppk) + (size_t)(4 * w) * 16 * 64 + l;
  HV bq[2][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) bq[0][t] = pb[(size_t)t * 16 * 64];
#pragma unroll 2
  for (int kt = 0; kt < 16; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < 16) {
#pragma unroll
      for (int t = 0; t < 4; ++t) bq[cur ^ 1][t] = pb[((size_t)t * 16 + kt + 1) * 64];
    }
    HV a[JT], al[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
      const int o = (jt * 16 + lr) * HLD + kt * 32 + 8 * g;
      a[jt] = *reinterpret_cast<const HV*>(&A0[o]);
      if (SPLIT) al[jt] = *reinterpret_cast<const HV*>(&A1[o]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int jt = 0; jt < JT; ++jt) {
        acc[t][jt] = mfma16(a[jt], bq[cur][t], acc[t][jt]);
        if (SPLIT) acc[t][jt] = mfma16(al[jt], bq[cur][t], acc[t][jt]);
      }
  }
#pragma unroll
  for (int jt = 0; jt < JT; ++jt)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int jr = jt * 16 + acc_row(g, q);
      const int b = j0 + jr;
      if (b < nb) {
        const float s = sc[jr];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          st_stream<NT>(&out[(size_t)b * BLEN + (4 * w + t) * 16 + lr], acc[t][jt][q] * s);
      }
    }
}

// One workgroup applies M^-1 to NBW consecutive blocks: every wave fills
// (phase 1), then every wave multiplies (phases 2 and 3).
template <int MODE, typename H, int NBW, bool SPLIT, bool NT = false>
__global__ __launch_bounds__(GT) void k_precond_h(int nb, const H* __restrict__ Ppk,
                                                  float p_unscale,
                                                  const float* __restrict__ invh,
                                                  KrylovVectors<float> v,
                                                  const float* __restrict__ pin,
                                                  float* __restrict__ pout, SolverState* st) {
  constexpr int NS = SPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(16))) H As[NS][NBW * HLD];
  __shared__ float sc[NBW];
  if (MODE != PREC_PLAIN && st->status != ST_RUNNING) return;
  const PrecScalars c = prec_scalars<MODE>(st);
  const int j0 = (int)blockIdx.x * NBW;
  float* __restrict__ out = MODE == PREC_PLAIN ? pout : (MODE == PREC_S ? v.z2 : v.z1);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  FillRegs<NBW, GW> R;
  precond_load<MODE, NBW, GW, NT>(nb, j0, w, l, v, pin, R);
  precond_convert<MODE, H, NBW, SPLIT, GW, NT>(nb, j0, w, l, p_unscale, invh, v, c, R, As[0],
                                           As[NS - 1], sc);
  __syncthreads();
  precond_gemm<H, NBW, SPLIT, NT>(nb, j0, w, l, Ppk, As[0], As[NS - 1], sc, out);
}

// Inner-solve start: r = rhat = (float)(r64 * scale), x = 0, ||r||^2, and the
// iteration-0 scalars (BiCGSTAB.cu:435-471 with x0 = 0).
__global__ __launch_bounds__(WG_THREADS) void k_inner_init(long long n, const double* __restrict__ r64,
                                                           double scale, KrylovVectors<float> v,
                                                           SolverState* st, double* part) {
  double s[1] = {0.0};
  const long long i0 = (long long)blockIdx.x * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    double rd[8];
    float r[8], z[8];
    load8(r64 + i0, rd);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      r[e] = (float)(rd[e] * scale);
      z[e] = 0.f;
      s[0] += (double)r[e] * (double)r[e];
    }
    store8(v.r + i0, r);
    store8(v.rhat + i0, r);
    store8(v.x[0] + i0, z);
  }
  grid_store_partial<1>(s, part);
}

__global__ __launch_bounds__(WG_THREADS) void k_inner_init_fin(SolverState* st, const double* part,
                                                               int nparts) {
  __shared__ SolverState S;
  double tot[1];
  fin_load<1>(st, S, part, nparts, tot);
  if (threadIdx.x == 0) {
    S.nu2 = 0.0;
    S.err = sqrt(tot[0]);
    S.err_init = S.err;
    S.err_opt = S.err;
    S.rhrh = tot[0];
    begin_iteration(&S, tot[0], tot[0]);
  }
  fin_store(st, S);
}

// Fold the FP32 increment into the FP64 iterate: x64 += scale * x32[k];
// x32[k] = 0, with k = xcur (zero = 1: the recurrence continues from this
// iterate) or xopt (zero = 0: the process ends here).
__global__ __launch_bounds__(WG_THREADS) void k_fold(long long n, double* __restrict__ x64,
                                                     KrylovVectors<float> v,
                                                     const SolverState* st, double scale,
                                                     int zero) {
  float* __restrict__ d = v.x[zero ? st->xcur : st->xopt];
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS) {
    x64[i] = fma(scale, (double)d[i], x64[i]);
    if (zero) d[i] = 0.f;
  }
}

// Residual replacement: r32 = r64 * inv_scale (the true FP64 residual of
// the folded iterate, in the scaling of the running Krylov process), with
// (rhat, r) and |r|^2 for the scalars.
__global__ __launch_bounds__(WG_THREADS) void k_replace(long long n, const double* __restrict__ r64,
                                                        double inv_scale, KrylovVectors<float> v,
                                                        double* part) {
  double s[2] = {0.0, 0.0};
  const long long i0 = (long long)blockIdx.x * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    double rd[8];
    float r[8], rh[8];
    load8(r64 + i0, rd);
    load8(v.rhat + i0, rh);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      r[e] = (float)(rd[e] * inv_scale);
      s[0] += (double)rh[e] * (double)r[e];
      s[1] += (double)r[e] * (double)r[e];
    }
    store8(v.r + i0, r);
  }
  grid_store_partial<2>(s, part);
}

// The scalars after a replacement. If the process had stopped (converged on
// its recurrence residual), the next iteration starts here; otherwise the
// iteration already begun takes the replaced residual's rho and |r|^2.
__global__ __launch_bounds__(WG_THREADS) void k_replace_fin(SolverState* st, const double* part,
                                                            int nparts) {
  __shared__ SolverState S;
  double tot[2];
  fin_load<2>(st, S, part, nparts, tot);
  if (threadIdx.x == 0) {
    S.err = sqrt(tot[1]);
    S.err_opt = S.err;
    S.xopt = S.xcur;
    if (S.status == ST_CONVERGED) {
      S.status = ST_RUNNING;
      S.rho_prev = S.rho_curr;
      begin_iteration(&S, tot[0], tot[1]);
    } else {
      if (S.restart_pending) {  // rhat = r in the next K1
        S.rho_curr = tot[1];
        S.rhrh = tot[1];
      } else {
        S.rho_curr = tot[0];
      }
      S.rr = tot[1];
      S.beta = (S.rho_curr / (S.rho_prev + S.eps)) * (S.alpha / (S.omega + S.eps));
    }
  }
  fin_store(st, S);
}

__global__ void k_copy64(long long n, const double* __restrict__ a, double* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    b[i] = a[i];
}

__global__ void k_cast64to32(long long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    b[i] = (float)a[i];
}
__global__ void k_cast32to64(long long n, const float* __restrict__ a, double* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    b[i] = (double)a[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

// Blocks per GEMM workgroup: P_inv (512 KiB in FP16) is streamed from L2 once
// per workgroup, so 32 blocks halve that traffic against 16; 32 is also the
// faster choice on grids too small to give every CU a workgroup (4440 blocks:
// 14.4 us against 15.9 us for K1).
static inline int gemm_width(int) { return 32; }

template <typename H, bool SPLIT, int NBW>
void launch_precond_hw(hipStream_t s, int mode, int nb, const H* P, float unscale,
                       const float* invh, const KrylovVectors<float>& v, const float* pin,
                       float* pout, SolverState* st) {
  const int nwg = (nb + NBW - 1) / NBW;
  // K1 and K3 stream their vectors nontemporally when they are large
  const bool nt = (size_t)nb * BLEN * sizeof(float) >= STREAM_MIN_BYTES;
  if (mode == PREC_P && nt)
    k_precond_h<PREC_P, H, NBW, SPLIT, true><<<nwg, GT, 0, s>>>(nb, P, unscale, invh, v, pin, pout, st);
  else if (mode == PREC_P)
    k_precond_h<PREC_P, H, NBW, SPLIT><<<nwg, GT, 0, s>>>(nb, P, unscale, invh, v, pin, pout, st);
  else if (mode == PREC_S && nt)
    k_precond_h<PREC_S, H, NBW, SPLIT, true><<<nwg, GT, 0, s>>>(nb, P, unscale, invh, v, pin, pout, st);
  else if (mode == PREC_S)
    k_precond_h<PREC_S, H, NBW, SPLIT><<<nwg, GT, 0, s>>>(nb, P, unscale, invh, v, pin, pout, st);
  else
    k_precond_h<PREC_PLAIN, H, NBW, SPLIT><<<nwg, GT, 0, s>>>(nb, P, unscale, invh, v, pin, pout, st);
}

template <typename H, bool SPLIT>
void launch_precond_hh(hipStream_t s, int mode, int nb, const H* P, float unscale,
                       const float* invh, const KrylovVectors<float>& v, const float* pin,
                       float* pout, SolverState* st) {
  if (gemm_width(nb) == 32)
    launch_precond_hw<H, SPLIT, 32>(s, mode, nb, P, unscale, invh, v, pin, pout, st);
  else
    launch_precond_hw<H, SPLIT, 16>(s, mode, nb, P, unscale, invh, v, pin, pout, st);
}

template <typename H>
void launch_precond_h(hipStream_t s, int mode, int nb, const void* Ppk, float unscale,
                      bool split, const float* invh, const KrylovVectors<float>& v,
                      const float* pin, float* pout, SolverState* st) {
  const H* P = static_cast<const H*>(Ppk);
  if (split) launch_precond_hh<H, true>(s, mode, nb, P, unscale, invh, v, pin, pout, st);
  else launch_precond_hh<H, false>(s, mode, nb, P, unscale, invh, v, pin, pout, st);
}

void launch_precond_mixed(const MixedWorkspace& M, hipStream_t s, int precision, int mode,
                          const DevGrid<float>& g, const float* pin, float* pout, SolverState* st) {
  if (M.use_dst) {
    KrylovVectors<float> v = M.v32;
    if (mode == PREC_PLAIN) {
      v.p = const_cast<float*>(pin);
      v.z1 = pout;
    }
    launch_precond_dst32(s, mode, g.nb, M.dst32, g.h, v, st);
    return;
  }
  if (precision == CUP3D_PREC_MIXED_FP16)
    launch_precond_h<_Float16>(s, mode, g.nb, M.d_P16, 1.0f / P16_SCALE, M.split, g.invh, M.v32,
                               pin, pout, st);
  else
    launch_precond_h<__bf16>(s, mode, g.nb, M.d_Pbf, 1.0f, M.split, g.invh, M.v32, pin, pout,
                             st);
}

template <int MODE>
void spmv32(hipStream_t s, const DevGrid<float>& g, const float* x, float* y, const float* a,
            SolverState* st, double* part, hipEvent_t mid = nullptr) {
  spmv_launch<float, MODE>(s, g, x, y, a, nullptr, nullptr, st, part, mid);
}

int grid_1d(long long n) { return (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 8192); }
int grid_vec(long long n) { return (int)((n + 4 * BLEN - 1) / (4 * BLEN)); }

SolverState inner_state(const cup3d_poisson_config& c, double tol_rel) {
  SolverState s;
  std::memset(&s, 0, sizeof(s));
  s.alpha = s.beta = s.omega = 1.0;
  s.eps = 1e-21;
  s.rho_prev = s.rho_curr = 1.0;
  s.err_init = s.err_opt = s.err = 1e50;
  s.tol_abs = 0.0;
  s.tol_rel = tol_rel;
  s.max_iter = c.max_iter;
  s.max_restarts = c.max_restarts;
  s.status = ST_RUNNING;
  s.p_zero = 1;
  return s;
}

#define MCHECK(expr)                                              \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) {                                       \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);    \
      return -1;                                                  \
    }                                                             \
  } while (0)

}  // namespace

int mixed_init_pinv(MixedWorkspace& M, const std::vector<double>& P, std::string& err) {
  // fragment (it, kt): lane l, element e = P[32 kt + 8 (l/16) + e][16 it + l%16]
  std::vector<_Float16> p16((size_t)BLEN * BLEN);
  std::vector<__bf16> pbf((size_t)BLEN * BLEN);
  for (int it = 0; it < 32; ++it)
    for (int kt = 0; kt < 16; ++kt)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          const double v = P[(size_t)(32 * kt + 8 * (l / 16) + e) * BLEN + 16 * it + l % 16];
          const size_t o = (((size_t)it * 16 + kt) * 64 + l) * 8 + e;
          const float vs = (float)(v * P16_SCALE);
          p16[o] = (_Float16)vs;
          pbf[o] = (__bf16)(float)v;
        }
  M.p16_scale = P16_SCALE;
  MCHECK(hipMalloc(&M.d_P16, p16.size() * sizeof(_Float16)));
  MCHECK(hipMalloc(&M.d_Pbf, pbf.size() * sizeof(__bf16)));
  MCHECK(hipMemcpy(M.d_P16, p16.data(), p16.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  MCHECK(hipMemcpy(M.d_Pbf, pbf.data(), pbf.size() * sizeof(__bf16), hipMemcpyHostToDevice));
  return 0;
}

int mixed_set_topology(MixedWorkspace& M, int nb, long long n, std::string& err) {
  // the buffers grow only: a mesh change that does not add cells keeps them
  if (n > M.cap_n || !M.d_vec32 || !M.d_res64) {
    if (M.d_vec32) (void)hipFree(M.d_vec32);
    if (M.d_res64) (void)hipFree(M.d_res64);
    M.d_vec32 = M.d_res64 = nullptr;
    M.cap_n = 0;
    MCHECK(hipMalloc(&M.d_vec32, (size_t)11 * n * sizeof(float)));
    MCHECK(hipMalloc(&M.d_res64, (size_t)2 * n * sizeof(double)));
    M.cap_n = n;
  }
  M.nb = nb;
  M.n = n;
  float* base = static_cast<float*>(M.d_vec32);
  KrylovVectors<float>& v = M.v32;
  for (int k = 0; k < 3; ++k) v.x[k] = base + k * n;
  v.r = base + 3 * n;
  v.rhat = base + 4 * n;
  v.p = base + 5 * n;
  v.nu = base + 6 * n;
  v.t = base + 7 * n;
  v.z1 = base + 8 * n;
  v.z2 = base + 9 * n;
  v.b = base + 10 * n;
  return 0;
}

int mixed_solve(MixedWorkspace& M, hipStream_t s, const cup3d_poisson_config& cfg,
                const DevGrid<double>& g64, const DevGrid<float>& g32,
                const KrylovVectors<double>& v64, SolverState* st, SolverState* h_st,
                double* part, int origin_block, const double* d_rhs, double* d_x,
                const double* d_x0, cup3d_solve_report* rep, std::string& err) {
  const long long n = M.n;
  const KrylovVectors<float>& v = M.v32;
  double* x64 = v64.x[0];
  double* b64 = v64.b;
  double* r64 = v64.r;
  // getVec + ||b|| (FP64)
  SolverState init = inner_state(cfg, 0.0);
  init.tol_abs = cfg.tol_abs;
  *h_st = init;
  MCHECK(hipMemcpyAsync(st, h_st, sizeof(SolverState), hipMemcpyHostToDevice, s));
  const long long zero_row =
      (cfg.bmean == 1 || cfg.bmean == 3) && origin_block >= 0 ? (long long)origin_block * BLEN : -1;
  launch_getvec64(s, n, zero_row, d_rhs, d_x0, b64, x64, st, part, cfg.tol_rhs);
  MCHECK(hipMemcpyAsync(h_st, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
  MCHECK(hipStreamSynchronize(s));
  const double bnorm2 = h_st->bnorm2;
  const double tol_abs = h_st->tol_abs;
  int total_iters = 0, total_restarts = 0, steps = 0, replacements = 0, status = ST_MAXITER;
  const int poll = cfg.poll_every > 0 ? cfg.poll_every : 8;
  auto true_resid = [&](double& out) -> int {  // |b - A x64| in FP64, r64 = b - A x64
    launch_spmv64(s, SPMV_RESID, g64, x64, nullptr, b64, r64, nullptr, st, part);
    MCHECK(hipMemcpyAsync(h_st, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    MCHECK(hipStreamSynchronize(s));
    out = std::sqrt(h_st->true_rr);
    return 0;
  };
  double rnorm = 0;
  if (true_resid(rnorm)) return -1;
  const double err_init = rnorm;
  const double tol = std::max(tol_abs, cfg.tol_rel * err_init);
  bool done = false;
  if (!std::isfinite(rnorm)) {
    status = ST_NONFINITE;
    done = true;
  } else if (rnorm <= tol) {
    status = ST_CONVERGED;
    done = true;
  }
  // One FP32 Krylov process per refinement step, on A d = r / |r|. The
  // process is not restarted when its recurrence residual has fallen by
  // REPLACE_DROP: the FP32 increment is folded into the FP64 iterate and the
  // recurrence residual replaced by the true FP64 residual (reliable updates,
  // as the residual replacement of PoissonSolverAMR.cpp keeps a long
  // recurrence accurate), so the Krylov space built so far is kept. A new
  // refinement step (restart) follows only a breakdown or the iteration cap.
#ifndef CUP3D_REPLACE_DROP
#define CUP3D_REPLACE_DROP 1e-1
#endif
  constexpr double REPLACE_DROP = CUP3D_REPLACE_DROP;  // 0: restart at every stop
#ifndef CUP3D_STALL_ITERS
#define CUP3D_STALL_ITERS 50
#endif
  // The window doubles after every stall restart that did not halve the FP64
  // residual: a recurrence that needs longer than the window for its first
  // halving (BiCGSTAB's residual often rises before it falls) would otherwise
  // repeat the same unproductive restart until max_refine is spent.
  int stall_iters = CUP3D_STALL_ITERS;
  while (!done && steps <= cfg.max_refine) {
    const double scale = rnorm;
    const double inner_rel = std::max(cfg.inner_tol, 0.5 * tol / scale);
    *h_st = inner_state(cfg, inner_rel);
    MCHECK(hipMemcpyAsync(st, h_st, sizeof(SolverState), hipMemcpyHostToDevice, s));
    k_inner_init<<<grid_vec(n), WG_THREADS, 0, s>>>(n, r64, 1.0 / scale, v, st, part);
    k_inner_init_fin<<<1, WG_THREADS, 0, s>>>(st, part, grid_vec(n));
    ++steps;
    double last = 1.0;  // recurrence residual (scaled) at the last replacement
    double progress = 1.0;  // residual at the last halving, and its iteration
    int progress_iter = 0;
    int guard = 0;
    while (true) {
      // timed iteration of the batch: one in the middle of it, away from the
      // idle GPU of the host poll; kept only if the batch got past it (the
      // kernels of iterations after convergence exit at once)
      const int tm_it = poll / 2;
      const int iter0 = h_st->iter;
      for (int it = 0; it < poll; ++it) {
        // event pairs (2k, 2k+1) bracket the main kernel of K(k+1)
        const bool tm = M.prof && it == tm_it;
        if (tm) MCHECK(hipEventRecord(M.pev[0], s));
        launch_precond_mixed(M, s, cfg.precision, PREC_P, g32, nullptr, nullptr, st);
        if (tm) MCHECK(hipEventRecord(M.pev[1], s));
        if (tm) MCHECK(hipEventRecord(M.pev[2], s));
        spmv32<SPMV_DOT1>(s, g32, v.z1, v.nu, v.rhat, st, part, tm ? M.pev[3] : nullptr);
        if (tm) MCHECK(hipEventRecord(M.pev[4], s));
        launch_precond_mixed(M, s, cfg.precision, PREC_S, g32, nullptr, nullptr, st);
        if (tm) MCHECK(hipEventRecord(M.pev[5], s));
        if (tm) MCHECK(hipEventRecord(M.pev[6], s));
        spmv32<SPMV_DOT2>(s, g32, v.z2, v.t, v.r, st, part, tm ? M.pev[7] : nullptr);
        if (tm) MCHECK(hipEventRecord(M.pev[8], s));
        launch_k5<float>(s, n, v, st, part, tm ? M.pev[9] : nullptr);
      }
      MCHECK(hipGetLastError());
      MCHECK(hipMemcpyAsync(h_st, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
      MCHECK(hipStreamSynchronize(s));
      if (M.prof && h_st->iter - iter0 > tm_it) {
        for (int k = 0; k < 5; ++k) {
          float ms = 0.f;
          MCHECK(hipEventElapsedTime(&ms, M.pev[2 * k], M.pev[2 * k + 1]));
          M.prof_ms[k] += ms;
        }
        ++M.prof_n;
      }
      if (++guard > cfg.max_iter / poll + 4 + 2 * replacements) {
        err = "inner solve did not stop";
        return -1;
      }
      int ist = h_st->status;
      if (h_st->err_opt < 0.5 * progress) {
        progress = h_st->err_opt;
        progress_iter = h_st->iter;
      }
      // a recurrence that has not halved its best residual for STALL_ITERS
      // iterations is restarted from the FP64 residual of its best iterate
      const bool stalled = ist == ST_RUNNING && h_st->iter - progress_iter > stall_iters;
      if (stalled) ist = ST_MAXITER;
      else if (ist == ST_RUNNING && h_st->err > REPLACE_DROP * last) continue;
      if (ist == ST_NONFINITE) {
        status = ST_NONFINITE;
        done = true;
        break;
      }
      if ((ist == ST_RUNNING || ist == ST_CONVERGED) && REPLACE_DROP > 0) {
        k_fold<<<grid_1d(n), WG_THREADS, 0, s>>>(n, x64, v, st, scale, 1);
        if (true_resid(rnorm)) return -1;
        ++replacements;
        if (!std::isfinite(rnorm)) {
          status = ST_NONFINITE;
          done = true;
          break;
        }
        if (rnorm <= tol) {
          status = ST_CONVERGED;
          done = true;
          break;
        }
        k_replace<<<grid_vec(n), WG_THREADS, 0, s>>>(n, r64, 1.0 / scale, v, part);
        k_replace_fin<<<1, WG_THREADS, 0, s>>>(st, part, grid_vec(n));
        last = rnorm / scale;
        progress = last;  // the stall window restarts at every replacement
        progress_iter = h_st->iter;
        continue;
      }
      // breakdown or iteration cap (or a stop without replacements): keep the
      // best iterate, restart
      k_fold<<<grid_1d(n), WG_THREADS, 0, s>>>(n, x64, v, st, scale, 0);
      if (true_resid(rnorm)) return -1;
      if (!std::isfinite(rnorm)) {
        status = ST_NONFINITE;
        done = true;
      } else if (rnorm <= tol) {
        status = ST_CONVERGED;
        done = true;
      } else if (stalled && rnorm > 0.5 * scale) {
        stall_iters *= 2;
      }
      break;
    }
    MCHECK(hipMemcpyAsync(h_st, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    MCHECK(hipStreamSynchronize(s));
    total_iters += h_st->iters_done;
    total_restarts += h_st->restarts;
  }
  k_copy64<<<grid_1d(n), WG_THREADS, 0, s>>>(n, x64, d_x);
  MCHECK(hipGetLastError());
  MCHECK(hipStreamSynchronize(s));
  if (rep) {
    rep->status = status;
    rep->iterations = total_iters;
    rep->restarts = total_restarts;
    rep->refinements = replacements;  // FP64 residual recomputations with a correction
    rep->err_init = err_init;
    rep->err_final = rnorm;
    rep->rhs_norm = std::sqrt(bnorm2);
    rep->true_rel_resid = bnorm2 > 0 ? rnorm / std::sqrt(bnorm2) : rnorm;
    rep->guess_resid = err_init;
  }
  return status;
}

int mixed_apply_operator32(MixedWorkspace& M, hipStream_t s, const DevGrid<float>& g32,
                           SolverState* st, double* part, const double* d_x, double* d_y,
                           std::string& err) {
  float* in = reinterpret_cast<float*>(M.d_res64);
  float* out = in + M.n;
  k_cast64to32<<<grid_1d(M.n), WG_THREADS, 0, s>>>(M.n, d_x, in);
  spmv32<SPMV_PLAIN>(s, g32, in, out, nullptr, st, part);
  k_cast32to64<<<grid_1d(M.n), WG_THREADS, 0, s>>>(M.n, out, d_y);
  MCHECK(hipGetLastError());
  return 0;
}

int mixed_apply_precond(MixedWorkspace& M, hipStream_t s, int precision, const DevGrid<float>& g32,
                        const double* d_p, double* d_z, std::string& err) {
  float* in = reinterpret_cast<float*>(M.d_res64);
  float* out = in + M.n;
  k_cast64to32<<<grid_1d(M.n), WG_THREADS, 0, s>>>(M.n, d_p, in);
  launch_precond_mixed(M, s, precision, PREC_PLAIN, g32, in, out, nullptr);
  k_cast32to64<<<grid_1d(M.n), WG_THREADS, 0, s>>>(M.n, out, d_z);
  MCHECK(hipGetLastError());
  return 0;
}

int mixed_time_kernel(MixedWorkspace& M, hipStream_t s, int which, int reps,
                      const DevGrid<float>& g32, SolverState* st, double* part, hipEvent_t e0,
                      hipEvent_t e1, double* avg_ms, double* bytes, double* flops,
                      std::string& err) {
  const KrylovVectors<float>& v = M.v32;
  const int precision = M.timing_precision;
  auto launch = [&]() {
    switch (which) {
      case 0: spmv32<SPMV_PLAIN>(s, g32, v.z1, v.nu, nullptr, st, part); break;
      case 1: launch_precond_mixed(M, s, precision, PREC_PLAIN, g32, v.p, v.z1, st); break;
      case 2:
        launch_k5<float>(s, M.n, v, st, part);
        break;
      case 3: launch_precond_mixed(M, s, precision, PREC_P, g32, nullptr, nullptr, st); break;
      default: spmv32<SPMV_DOT1>(s, g32, v.z1, v.nu, v.rhat, st, part); break;
    }
  };
  launch();
  MCHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  MCHECK(hipEventRecord(e1, s));
  MCHECK(hipEventSynchronize(e1));
  float ms = 0;
  MCHECK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / reps;
  const double N = (double)M.n;
  // algorithmic FP32 words per cell and flops per cell; a GEMM pass is
  // 2 x 512 flops per cell, two passes with the split input
  const double words[5] = {2, 2, 8, 5, 3};
  const double passes = M.split ? 2.0 : 1.0;
  const double pf = M.use_dst ? 96.0 : 1024 * passes;  // sine transforms: 96 flops per cell
  const double fpc[5] = {8, pf, 8, pf + 4, 10};
  *bytes = 4.0 * words[which] * N;
  *flops = fpc[which] * N;
  return 0;
}

void mixed_free(MixedWorkspace& M) {
  if (M.d_P16) (void)hipFree(M.d_P16);
  if (M.d_Pbf) (void)hipFree(M.d_Pbf);
  if (M.d_vec32) (void)hipFree(M.d_vec32);
  if (M.d_res64) (void)hipFree(M.d_res64);
  M.d_P16 = M.d_Pbf = M.d_vec32 = M.d_res64 = nullptr;
  for (hipEvent_t& e : M.pev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  M.prof = false;
}

}  // namespace cup3d
