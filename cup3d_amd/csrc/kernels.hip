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

namespace cup3d {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Scalar state machine (single thread of a last workgroup).
// ---------------------------------------------------------------------------

// Start of iteration k: the breakdown test, beta and the restart of
// BiCGSTAB.cu:477-515, and the iteration cap of :474-475.
__device__ void begin_iteration(SolverState* st, double rho, double rr) {
  if (st->iter >= st->max_iter) {
    st->status = ST_MAXITER;
    return;
  }
  st->rho_curr = rho;
  st->rr = rr;
  const bool serious = rho * rho < 1e-16 * rr * st->rhrh;
  st->beta = (st->rho_curr / (st->rho_prev + st->eps)) * (st->alpha / (st->omega + st->eps));
  if (serious && st->max_restarts > 0) {
    st->restarts += 1;
    if (st->restarts >= st->max_restarts) {
      st->status = ST_BREAKDOWN;
      return;
    }
    st->restart_pending = 1;  // rhat = r (done by K1)
    st->p_zero = 1;           // p = nu = 0
    st->rho_curr = rr;        // ||rhat||^2 after rhat = r
    st->rhrh = rr;
    st->rho_prev = 1.0;       // breakdown_update (:263-269)
    st->alpha = 1.0;
    st->omega = 1.0;
    st->beta = (st->rho_curr / (st->rho_prev + st->eps)) * (st->alpha / (st->omega + st->eps));
  }
  st->iter += 1;
}

// ---------------------------------------------------------------------------
// getVec (PoissonSolverExp.cpp:375-418): b = rhs, x = x0, both zeroed at the
// bMean cell for bMeanConstraint 1 and 3; also ||b||^2.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(WG_THREADS) void k_getvec(long long n, long long zero_row,
                                                       const double* __restrict__ rhs,
                                                       const double* __restrict__ x0,
                                                       T* __restrict__ b, T* __restrict__ x,
                                                       SolverState* st, double* part,
                                                       double tol_b) {
  double v[1] = {0.0};
  const long long base = (long long)xcd_remap(blockIdx.x, gridDim.x) * (4 * BLEN);
  for (int e = threadIdx.x; e < 4 * BLEN; e += WG_THREADS) {
    const long long i = base + e;
    if (i < n) {
      double bv = rhs[i], xv = x0 ? x0[i] : 0.0;
      if (i == zero_row) bv = xv = 0.0;
      b[i] = (T)bv;
      x[i] = (T)xv;
      v[0] += bv * bv;
    }
  }
  double tot[1];
  if (grid_reduce_last<1>(v, part, &st->counters[0], tot) && threadIdx.x == 0) {
    st->bnorm2 = tot[0];
    const double tb = tol_b * sqrt(tot[0]);
    if (tb > st->tol_abs) st->tol_abs = tb;
  }
}

// ---------------------------------------------------------------------------
// Matrix-free operator y = A x (PoissonSolverExp.cpp:305-367 + interface rows)
// with the bMean handling of hd_cusparseSpMV (BiCGSTAB.cu:392-413) and a
// fused reduction epilogue.
//
// One wavefront per block: lane = (ix, iy), each lane walks its z column in
// registers; the block is staged in LDS for the x/y neighbours; face halos
// of same-level neighbours are read straight from HBM (the z faces as
// coalesced 512-byte rows).
// ---------------------------------------------------------------------------
template <typename T, int MODE, int BM>
__global__ __launch_bounds__(WG_THREADS) void k_spmv(DevGrid<T> g, const T* __restrict__ x,
                                                     T* __restrict__ y, const T* __restrict__ a,
                                                     T* __restrict__ c, T* __restrict__ c2,
                                                     SolverState* st, double* part) {
  __shared__ T xs[4][BLEN];
  __shared__ T ys[4][BLEN];
  if (MODE != SPMV_PLAIN && MODE != SPMV_RESID && st->status != ST_RUNNING) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ix = lane & 7, iy = lane >> 3;
  const int b = xcd_remap(blockIdx.x, gridDim.x) * 4 + w;
  const bool valid = b < g.nb;
  const size_t off = (size_t)b * BLEN;
  T col[BS];
  if (valid) {
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      col[iz] = x[off + iz * 64 + lane];
      xs[w][iz * 64 + lane] = col[iz];
    }
  }
  __syncthreads();
  double msum = 0.0;
  if (valid) {
    const int* nb6 = g.nbr + b * 6;
    const int nxm = nb6[0], nxp = nb6[1], nym = nb6[2], nyp = nb6[3], nzm = nb6[4], nzp = nb6[5];
    const T h = g.h[b];
    const T zlo = nzm >= 0 ? x[(size_t)nzm * BLEN + 7 * 64 + lane] : col[0];
    const T zhi = nzp >= 0 ? x[(size_t)nzp * BLEN + lane] : col[BS - 1];
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      const T c0 = col[iz];
      const int li = iz * 64 + lane;
      T xm, xp, ym, yp;
      if (ix > 0) xm = xs[w][li - 1];
      else xm = nxm >= 0 ? x[(size_t)nxm * BLEN + iz * 64 + iy * 8 + 7] : c0;
      if (ix < 7) xp = xs[w][li + 1];
      else xp = nxp >= 0 ? x[(size_t)nxp * BLEN + iz * 64 + iy * 8] : c0;
      if (iy > 0) ym = xs[w][li - 8];
      else ym = nym >= 0 ? x[(size_t)nym * BLEN + iz * 64 + 56 + ix] : c0;
      if (iy < 7) yp = xs[w][li + 8];
      else yp = nyp >= 0 ? x[(size_t)nyp * BLEN + iz * 64 + ix] : c0;
      const T zm = iz > 0 ? col[iz - 1] : zlo;
      const T zp = iz < BS - 1 ? col[iz + 1] : zhi;
      ys[w][li] = h * ((((xm + xp) + (ym + yp)) + (zm + zp)) - (T)6 * c0);
      msum += (double)c0;
    }
  }
  __syncthreads();
  if (valid) {
    const int r0 = g.irow_ptr[b], r1 = g.irow_ptr[b + 1];
    for (int r = r0 + lane; r < r1; r += 64) {
      T acc = 0;
      for (int e = g.irow_ent[r]; e < g.irow_ent[r + 1]; ++e) acc += g.ent_val[e] * x[g.ent_col[e]];
      ys[w][g.irow_cell[r]] += acc;
    }
  }
  __syncthreads();
  const long long row = g.bmean_row;  // origin cell for BM 1 and 3, else -1
  const bool origin = valid && row == (long long)off;
  T yv[BS];
  if (valid) {
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) yv[iz] = ys[w][iz * 64 + lane];
    if (BM == 3 && origin && lane == 0) yv[0] = g.h[b] * col[0];  // row h*I (.cpp:311-318)
  }
  // epilogue: stores and partial sums (the bMean row is left to the last WG)
  double v[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const double h = (double)g.h[b];
    v[0] = h * h * h * msum;
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      const size_t i = off + iz * 64 + lane;
      const bool skip = (BM == 1) && origin && lane == 0 && iz == 0;
      const T yy = yv[iz];
      if (MODE == SPMV_PLAIN) {
        y[i] = yy;
      } else if (MODE == SPMV_INIT) {
        y[i] = yy;
        const T rr = a[i] - yy;
        c[i] = rr;
        c2[i] = rr;
        if (!skip) {
          v[1] += (double)yy * (double)yy;
          v[2] += (double)rr * (double)rr;
        }
      } else if (MODE == SPMV_DOT1) {
        y[i] = yy;
        if (!skip) v[1] += (double)a[i] * (double)yy;
      } else if (MODE == SPMV_DOT2) {
        y[i] = yy;
        if (!skip) {
          v[1] += (double)a[i] * (double)yy;
          v[2] += (double)yy * (double)yy;
        }
      } else if (MODE == SPMV_RESID) {
        const T rr = a[i] - yy;
        c[i] = rr;
        if (!skip) v[1] += (double)rr * (double)rr;
      }
    }
  }
  double tot[3];
  if (!grid_reduce_last<3>(v, part, &st->counters[1], tot)) return;
  if (threadIdx.x != 0) return;
  // bMean = 1: y[row] = sum_b h_b^3 sum x  (BiCGSTAB.cu:406-407)
  if (BM == 1 && row >= 0) {
    const T yr = (T)tot[0];
    if (MODE == SPMV_PLAIN) {
      y[row] = yr;
    } else if (MODE == SPMV_INIT) {
      y[row] = yr;
      const T rr = a[row] - yr;
      c[row] = rr;
      c2[row] = rr;
      tot[1] += (double)yr * (double)yr;
      tot[2] += (double)rr * (double)rr;
    } else if (MODE == SPMV_DOT1) {
      y[row] = yr;
      tot[1] += (double)a[row] * (double)yr;
    } else if (MODE == SPMV_DOT2) {
      y[row] = yr;
      tot[1] += (double)a[row] * (double)yr;
      tot[2] += (double)yr * (double)yr;
    } else if (MODE == SPMV_RESID) {
      const T rr = a[row] - yr;
      c[row] = rr;
      tot[1] += (double)rr * (double)rr;
    }
  }
  st->mean_sum = tot[0];
  if (MODE == SPMV_INIT) {
    // initial residual (BiCGSTAB.cu:435-471)
    st->nu2 = tot[1];
    st->err = sqrt(tot[2]);
    st->err_init = st->err;
    st->err_opt = st->err;
    st->rhrh = tot[2];
    begin_iteration(st, tot[2], tot[2]);
  } else if (MODE == SPMV_DOT1) {
    st->dot_a = tot[1];
    st->alpha = st->rho_curr / (tot[1] + st->eps);  // set_alpha (:276-279)
    st->p_zero = 0;
    st->restart_pending = 0;
  } else if (MODE == SPMV_DOT2) {
    st->dot_ts = tot[1];
    st->dot_tt = tot[2];
    st->omega = tot[1] / (tot[2] + st->eps);  // set_omega (:281-284)
  } else if (MODE == SPMV_RESID) {
    st->true_rr = tot[1];
  }
}

// ---------------------------------------------------------------------------
// K5: x' = x + alpha z1 + omega z2 (into a free rotating buffer), r = s - omega t,
// then rho' = (rhat, r), |r|, x_opt, convergence and the next iteration's
// scalars (BiCGSTAB.cu:543-607).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int third_buffer(int a, int b) {
  for (int k = 0; k < 3; ++k)
    if (k != a && k != b) return k;
  return 0;
}

template <typename T>
__global__ __launch_bounds__(WG_THREADS) void k_update(long long n, KrylovVectors<T> v,
                                                       SolverState* st, double* part) {
  if (st->status != ST_RUNNING) return;
  const int xc = st->xcur, xo = st->xopt, xn = third_buffer(xc, xo);
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  const T* __restrict__ xin = v.x[xc];
  T* __restrict__ xout = v.x[xn];
  double s[2] = {0.0, 0.0};
  const long long base = (long long)xcd_remap(blockIdx.x, gridDim.x) * (4 * BLEN);
  for (int e = threadIdx.x; e < 4 * BLEN; e += WG_THREADS) {
    const long long i = base + e;
    if (i < n) {
      xout[i] = fma(omega, v.z2[i], fma(alpha, v.z1[i], xin[i]));
      const T r = fma(-omega, v.t[i], v.r[i]);
      v.r[i] = r;
      s[0] += (double)v.rhat[i] * (double)r;
      s[1] += (double)r * (double)r;
    }
  }
  double tot[2];
  if (!grid_reduce_last<2>(s, part, &st->counters[2], tot) || threadIdx.x != 0) return;
  const double err = sqrt(tot[1]);
  st->err = err;
  st->iters_done = st->iter;
  st->xcur = xn;
  if (!(err == err) || err > 1.7e308) {
    st->status = ST_NONFINITE;
    return;
  }
  if (err < st->err_opt) {
    st->err_opt = err;
    st->xopt = xn;
    if (err <= st->tol_abs || err / st->err_init <= st->tol_rel) {
      st->status = ST_CONVERGED;
      return;
    }
  }
  st->rho_prev = st->rho_curr;  // set_rho (:607)
  begin_iteration(st, tot[0], tot[1]);
}

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
  double bq[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) bq[t] = pb[(size_t)t * 128 * 64];
  for (int kt = 0; kt < 128; ++kt) {
    const double av = Vs[lr * PC_LD + 4 * kt + lk];
    double bn[8];
    const int kn = kt + 1 < 128 ? kt + 1 : kt;
#pragma unroll
    for (int t = 0; t < 8; ++t) bn[t] = pb[((size_t)t * 128 + kn) * 64];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bq[t], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 8; ++t) bq[t] = bn[t];
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

// Copy the best iterate out: x_user = X[xopt] (PoissonSolverExp.cpp:452-472
// scatter; BiCGSTAB.cu:625 copies x_opt).
template <typename T>
__global__ __launch_bounds__(WG_THREADS) void k_copy_opt(long long n, KrylovVectors<T> v,
                                                         const SolverState* st,
                                                         double* __restrict__ out) {
  const T* __restrict__ src = v.x[st->xopt];
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    out[i] = (double)src[i];
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
}

template <int MODE, int BM>
static void spmv_bm(hipStream_t s, const DevGrid<double>& g, const double* x, double* y,
                    const double* a, double* c, double* c2, SolverState* st, double* part) {
  k_spmv<double, MODE, BM><<<wg_for_blocks(g.nb, 4), WG_THREADS, 0, s>>>(g, x, y, a, c, c2, st,
                                                                         part);
}

template <int MODE>
static void spmv_mode(hipStream_t s, const DevGrid<double>& g, const double* x, double* y,
                      const double* a, double* c, double* c2, SolverState* st, double* part) {
  switch (g.bmean) {
    case 1: spmv_bm<MODE, 1>(s, g, x, y, a, c, c2, st, part); break;
    case 3: spmv_bm<MODE, 3>(s, g, x, y, a, c, c2, st, part); break;
    default: spmv_bm<MODE, 0>(s, g, x, y, a, c, c2, st, part); break;
  }
}

void launch_spmv64(hipStream_t s, int mode, const DevGrid<double>& g, const double* x, double* y,
                   const double* a, double* c, double* c2, SolverState* st, double* part) {
  switch (mode) {
    case SPMV_PLAIN: spmv_mode<SPMV_PLAIN>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_INIT: spmv_mode<SPMV_INIT>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_DOT1: spmv_mode<SPMV_DOT1>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_DOT2: spmv_mode<SPMV_DOT2>(s, g, x, y, a, c, c2, st, part); break;
    case SPMV_RESID: spmv_mode<SPMV_RESID>(s, g, x, y, a, c, c2, st, part); break;
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
