// HIP kernels of the pressure-Poisson solve for gfx950 (MI355X).
//
// Vectors live in HBM as contiguous [nblocks][8][8][8] arrays, blockID-major,
// [iz][iy][ix] inside a block (Cubism's ScalarBlock order, the flat index
// j = blockID*512 + iz*64 + iy*8 + ix of PoissonSolverAMR.cpp:76).
//
// One classical BiCGSTAB iteration (BiCGSTAB.cu:474-608) runs as five fused
// kernels; every reduction is a deterministic two-pass sum (per-workgroup
// partials, then the last workgroup adds them in workgroup order) and the
// scalar recurrences run on the device, so the host only polls a status word
// every few iterations:
//   K1 k_precond<P>   p = r + beta (p - omega nu)  [p = r after a restart,
//                     rhat = r]; z1 = M^-1 p
//   K2 k_spmv<DOT1>   nu = A z1; alpha = rho / ((rhat, nu) + eps)
//   K3 k_precond<S>   s = r - alpha nu (in r); z2 = M^-1 s
//   K4 k_spmv<DOT2>   t = A z2; omega = (t, s) / ((t, t) + eps)
//   K5 k_update       x' = x + alpha z1 + omega z2; r = s - omega t;
//                     rho' = (rhat, r), |r|; x_opt / convergence / restart
// M^-1 = blockdiag(h_b^-1 P_inv) is one GEMM over all blocks on the FP64
// matrix cores (v_mfma_f64_16x16x4f64).

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"
#include "krylov.h"

namespace cup3d {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// K1 / K3: fused vector update + block-Jacobi GEMM on the FP64 matrix cores.
//
// A workgroup owns 16 blocks. Their updated vectors (p or s) are written to
// HBM and staged in LDS as 16 rows of 512 doubles (row stride 514, which
// keeps the fragment reads free of bank conflicts). Each wavefront computes
// 128 output cells x 16 blocks as 8 tiles of v_mfma_f64_16x16x4f64:
//   D[j][i] += sum_k V[j][k] * P[k][i]     (j: block, i: output cell)
// with the A operand (V) read from LDS and the B operand (P_inv) streamed
// from L2 in a fragment-packed layout: fragment (it, kt) holds
// P[4 kt + l/16][16 it + l%16] in lane l, 64 consecutive doubles.
// The reference computes z_b = invh_b * P_inv p_b with cublasDgemm(OP_T, OP_N)
// + blockDscal (BiCGSTAB.cu:526-528,553-555); P_inv is symmetric.
// ---------------------------------------------------------------------------
constexpr int PC_NB = 16;       // blocks per workgroup
constexpr int PC_LD = BLEN + 2; // LDS row stride (doubles)

template <int MODE>
__global__ __launch_bounds__(WG_THREADS) void k_precond_f64(int nb, const double* __restrict__ Ppk,
                                                            const double* __restrict__ invh,
                                                            KrylovVectors<double> v,
                                                            SolverState* st) {
  __shared__ double Vs[PC_NB * PC_LD];
  if (MODE != PREC_PLAIN && st->status != ST_RUNNING) return;
  const int j0 = xcd_remap(blockIdx.x, gridDim.x) * PC_NB;
  const double beta = MODE == PREC_P ? st->beta : 0.0;
  const double omega = MODE == PREC_P ? st->omega : 0.0;
  const double alpha = MODE == PREC_S ? st->alpha : 0.0;
  const int pzero = MODE == PREC_P ? st->p_zero : 0;
  const int restart = MODE == PREC_P ? st->restart_pending : 0;
  double* __restrict__ out = MODE == PREC_S ? v.z2 : v.z1;
  for (int e = threadIdx.x; e < PC_NB * BLEN; e += WG_THREADS) {
    const int j = e >> 9, k = e & (BLEN - 1);
    const int b = j0 + j;
    double val = 0.0;
    if (b < nb) {
      const size_t i = (size_t)b * BLEN + k;
      if (MODE == PREC_P) {
        // p = r + beta (p - omega nu)   (BiCGSTAB.cu:518-522)
        const double r = v.r[i];
        val = pzero ? r : fma(beta, fma(-omega, v.nu[i], v.p[i]), r);
        v.p[i] = val;
        if (restart) v.rhat[i] = r;
      } else if (MODE == PREC_S) {
        // s = r - alpha nu   (BiCGSTAB.cu:546-549)
        val = fma(-alpha, v.nu[i], v.r[i]);
        v.r[i] = val;
      } else {
        val = v.p[i];
      }
    }
    Vs[j * PC_LD + k] = val;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
  double4_t acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const double* __restrict__ pb = Ppk + (size_t)(8 * w) * 128 * 64 + l;
  // P fragments stream from L2 through a 4-stage register ring: the loads for
  // k-step kt+4 are issued while the MFMAs of kt run, hiding L2 latency.
  constexpr int NST = 4;
  double bq[NST][8];
#pragma unroll
  for (int st_ = 0; st_ < NST; ++st_)
#pragma unroll
    for (int t = 0; t < 8; ++t) bq[st_][t] = pb[((size_t)t * 128 + st_) * 64];
  for (int k0 = 0; k0 < 128; k0 += NST) {
#pragma unroll
    for (int st_ = 0; st_ < NST; ++st_) {
      const int kt = k0 + st_;
      const double av = Vs[lr * PC_LD + 4 * kt + lk];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[st_][t], acc[t], 0, 0, 0);
      const int kn = kt + NST < 128 ? kt + NST : kt;
#pragma unroll
      for (int t = 0; t < 8; ++t) bq[st_][t] = pb[((size_t)t * 128 + kn) * 64];
    }
  }
  // D layout of v_mfma_f64_16x16x4f64: register q of lane l holds
  // D[l/16 + 4q][l%16]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = j0 + lk + 4 * q;
    if (b < nb) {
      const double s = invh[b];
#pragma unroll
      for (int t = 0; t < 8; ++t) out[(size_t)b * BLEN + (8 * w + t) * 16 + lr] = acc[t][q] * s;
    }
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline int wg_for_blocks(int nb, int per) { return (nb + per - 1) / per; }

void launch_getvec64(hipStream_t s, long long n, long long zero_row, const double* rhs,
                     const double* x0, double* b, double* x, SolverState* st, double* part,
                     double tol_b) {
  const int nwg = (int)((n + 4 * BLEN - 1) / (4 * BLEN));
  k_getvec<double><<<nwg, WG_THREADS, 0, s>>>(n, zero_row, rhs, x0, b, x, st, part, tol_b);
  k_getvec_fin<<<1, WG_THREADS, 0, s>>>(st, part, nwg, tol_b);
}

void launch_spmv64(hipStream_t s, int mode, const DevGrid<double>& g, const double* x, double* y,
                   const double* a, double* c, double* c2, SolverState* st, double* part) {
  switch (mode) {
    case SPMV_PLAIN: spmv_launch<double, SPMV_PLAIN>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_INIT: spmv_launch<double, SPMV_INIT>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_DOT1: spmv_launch<double, SPMV_DOT1>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_DOT2: spmv_launch<double, SPMV_DOT2>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_RESID: spmv_launch<double, SPMV_RESID>(s, g, x, y, a, c, c2, st, part); break;
  }
}

void launch_precond64(hipStream_t s, int mode, int nb, const double* Ppk, const double* invh,
                      const KrylovVectors<double>& v, SolverState* st) {
  const int nwg = wg_for_blocks(nb, PC_NB);
  switch (mode) {
    case PREC_P: k_precond_f64<PREC_P><<<nwg, WG_THREADS, 0, s>>>(nb, Ppk, invh, v, st); break;
    case PREC_S: k_precond_f64<PREC_S><<<nwg, WG_THREADS, 0, s>>>(nb, Ppk, invh, v, st); break;
    default: k_precond_f64<PREC_PLAIN><<<nwg, WG_THREADS, 0, s>>>(nb, Ppk, invh, v, st); break;
  }
}

void launch_update64(hipStream_t s, long long n, const KrylovVectors<double>& v, SolverState* st,
                     double* part) {
  const int nwg = (int)((n + 4 * BLEN - 1) / (4 * BLEN));
  k_update<double><<<nwg, WG_THREADS, 0, s>>>(n, v, st, part);
  k_update_fin<<<1, WG_THREADS, 0, s>>>(st, part, nwg);
}

void launch_copy_opt64(hipStream_t s, long long n, const KrylovVectors<double>& v,
                       const SolverState* st, double* out) {
  const int nwg = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 4096);
  k_copy_opt<double><<<nwg, WG_THREADS, 0, s>>>(n, v, st, out);
}

// Fragment-packed P for the B operand: Ppk[(it*128 + kt)*64 + l] = P[4kt + l/16][16it + l%16].
void pack_pinv_f64(const std::vector<double>& P, std::vector<double>& pk) {
  pk.assign((size_t)BLEN * BLEN, 0.0);
  for (int it = 0; it < 32; ++it)
    for (int kt = 0; kt < 128; ++kt)
      for (int l = 0; l < 64; ++l)
        pk[((size_t)it * 128 + kt) * 64 + l] = P[(size_t)(4 * kt + l / 16) * BLEN + 16 * it + l % 16];
}

}  // namespace cup3d
