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
// A workgroup of 8 wavefronts owns 32 blocks (one workgroup per CU at 8192
// blocks). Their updated vectors (p or s) are written to HBM and staged in
// LDS as 32 rows of 512 doubles (row stride 514, which keeps the fragment
// reads free of bank conflicts). Each wavefront computes 64 output cells x
// 32 blocks as 4 x 2 tiles of v_mfma_f64_16x16x4f64:
//   D[j][i] += sum_k V[j][k] * P[k][i]     (j: block, i: output cell)
// with the A operand (V) read from LDS and the B operand (P_inv) streamed
// from L2 in a fragment-packed layout: fragment (it, kt) holds
// P[4 kt + l/16][16 it + l%16] in lane l, 64 consecutive doubles. P_inv
// (2 MiB) is streamed once per 32 blocks.
// The reference computes z_b = invh_b * P_inv p_b with cublasDgemm(OP_T, OP_N)
// + blockDscal (BiCGSTAB.cu:526-528,553-555); P_inv is symmetric.
// ---------------------------------------------------------------------------
constexpr int PC_W = 8;          // wavefronts per workgroup
constexpr int PC_T = PC_W * 64;  // threads per workgroup
constexpr int PC_LD = BLEN + 2;  // LDS row stride (doubles)

typedef double double2_t __attribute__((ext_vector_type(2)));

template <int MODE, int PC_NB>
__global__ __launch_bounds__(PC_T) void k_precond_f64(int nb, const double* __restrict__ Ppk,
                                                      const double* __restrict__ invh,
                                                      KrylovVectors<double> v, SolverState* st) {
  constexpr int JT = PC_NB / 16;  // block tiles
  __shared__ double Vs[PC_NB * PC_LD];
  if (MODE != PREC_PLAIN && st->status != ST_RUNNING) return;
  const int j0 = (int)blockIdx.x * PC_NB;
  const double beta = MODE == PREC_P ? st->beta : 0.0;
  const double omega = MODE == PREC_P ? st->omega : 0.0;
  const double alpha = MODE == PREC_S ? st->alpha : 0.0;
  const int pzero = MODE == PREC_P ? st->p_zero : 0;
  const int restart = MODE == PREC_P ? st->restart_pending : 0;
  double* __restrict__ out = MODE == PREC_S ? v.z2 : v.z1;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  // phase 1: wave w owns blocks [NPB w, NPB w + NPB), two at a time; element
  // u of a lane covers cells 128 (u % 4) + 2 l, +1 of block 2 h + u / 4. All
  // loads of a pair are issued before the first use.
  constexpr int NPB = PC_NB / PC_W;  // 4 or 2
#pragma unroll
  for (int hh = 0; hh < NPB / 2; ++hh) {
    double2_t ra[8], rb[8], rc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = j0 + NPB * w + 2 * hh + u / 4;
      const size_t i2 = ((size_t)b * BLEN + 128 * (u & 3) + 2 * l) / 2;
      if (b < nb) {
        if (MODE == PREC_P) {
          ra[u] = reinterpret_cast<const double2_t*>(v.r)[i2];
          rb[u] = reinterpret_cast<const double2_t*>(v.nu)[i2];
          rc[u] = reinterpret_cast<const double2_t*>(v.p)[i2];
        } else if (MODE == PREC_S) {
          ra[u] = reinterpret_cast<const double2_t*>(v.r)[i2];
          rb[u] = reinterpret_cast<const double2_t*>(v.nu)[i2];
        } else {
          ra[u] = reinterpret_cast<const double2_t*>(v.p)[i2];
        }
      } else {
        ra[u] = rb[u] = rc[u] = (double2_t){0.0, 0.0};
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int jj = NPB * w + 2 * hh + u / 4;
      const int b = j0 + jj;
      const size_t i2 = ((size_t)b * BLEN + 128 * (u & 3) + 2 * l) / 2;
      double2_t val = ra[u];
      if (MODE == PREC_P) {
        // p = r + beta (p - omega nu)   (BiCGSTAB.cu:518-522)
        if (!pzero) {
          val.x = fma(beta, fma(-omega, rb[u].x, rc[u].x), ra[u].x);
          val.y = fma(beta, fma(-omega, rb[u].y, rc[u].y), ra[u].y);
        }
        if (b < nb) {
          reinterpret_cast<double2_t*>(v.p)[i2] = val;
          if (restart) reinterpret_cast<double2_t*>(v.rhat)[i2] = ra[u];
        }
      } else if (MODE == PREC_S) {
        // s = r - alpha nu   (BiCGSTAB.cu:546-549)
        val.x = fma(-alpha, rb[u].x, ra[u].x);
        val.y = fma(-alpha, rb[u].y, ra[u].y);
        if (b < nb) reinterpret_cast<double2_t*>(v.r)[i2] = val;
      }
      const int k = 128 * (u & 3) + 2 * l;
      Vs[jj * PC_LD + k] = val.x;
      Vs[jj * PC_LD + k + 1] = val.y;
    }
  }
  __syncthreads();
  // phase 2: wave w owns output cells [64 w, 64 w + 64) of all PC_NB blocks
  const int lr = l & 15, lk = l >> 4;
  double4_t acc[4][JT];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) acc[t][jt] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const double* __restrict__ pb = Ppk + (size_t)(4 * w) * 128 * 64 + l;
  // P fragments stream from L2 through a 4-stage register ring: the loads for
  // k-step kt+4 are issued while the MFMAs of kt run, hiding L2 latency.
  constexpr int NST = 4;
  double bq[NST][4];
#pragma unroll
  for (int s_ = 0; s_ < NST; ++s_)
#pragma unroll
    for (int t = 0; t < 4; ++t) bq[s_][t] = pb[((size_t)t * 128 + s_) * 64];
  for (int k0 = 0; k0 < 128; k0 += NST) {
#pragma unroll
    for (int s_ = 0; s_ < NST; ++s_) {
      const int kt = k0 + s_;
      double a[JT];
#pragma unroll
      for (int jt = 0; jt < JT; ++jt) a[jt] = Vs[(16 * jt + lr) * PC_LD + 4 * kt + lk];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
          acc[t][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[jt], bq[s_][t], acc[t][jt], 0, 0, 0);
      const int kn = kt + NST < 128 ? kt + NST : kt;
#pragma unroll
      for (int t = 0; t < 4; ++t) bq[s_][t] = pb[((size_t)t * 128 + kn) * 64];
    }
  }
  // D layout of v_mfma_f64_16x16x4f64: register q of lane l holds
  // D[l/16 + 4q][l%16]
#pragma unroll
  for (int jt = 0; jt < JT; ++jt)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int b = j0 + jt * 16 + lk + 4 * q;
      if (b < nb) {
        const double s = invh[b];
#pragma unroll
        for (int t = 0; t < 4; ++t)
          out[(size_t)b * BLEN + (4 * w + t) * 16 + lr] = acc[t][jt][q] * s;
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

// Blocks per workgroup: 32 halves the L2 stream of P_inv but needs 8192
// blocks to give every CU a workgroup; smaller grids use 16.
static inline int precond64_width(int nb) { return nb >= 32 * 256 ? 32 : 16; }

template <int NB>
static void launch_precond64_w(hipStream_t s, int mode, int nb, const double* Ppk,
                               const double* invh, const KrylovVectors<double>& v,
                               SolverState* st) {
  const int nwg = wg_for_blocks(nb, NB);
  switch (mode) {
    case PREC_P: k_precond_f64<PREC_P, NB><<<nwg, PC_T, 0, s>>>(nb, Ppk, invh, v, st); break;
    case PREC_S: k_precond_f64<PREC_S, NB><<<nwg, PC_T, 0, s>>>(nb, Ppk, invh, v, st); break;
    default: k_precond_f64<PREC_PLAIN, NB><<<nwg, PC_T, 0, s>>>(nb, Ppk, invh, v, st); break;
  }
}

void launch_precond64(hipStream_t s, int mode, int nb, const double* Ppk, const double* invh,
                      const KrylovVectors<double>& v, SolverState* st) {
  if (precond64_width(nb) == 32) launch_precond64_w<32>(s, mode, nb, Ppk, invh, v, st);
  else launch_precond64_w<16>(s, mode, nb, Ppk, invh, v, st);
}

void launch_update64(hipStream_t s, long long n, const KrylovVectors<double>& v, SolverState* st,
                     double* part) {
  launch_k5<double>(s, n, v, st, part);
}

void launch_copy_opt64(hipStream_t s, long long n, const KrylovVectors<double>& v,
                       const SolverState* st, double* out) {
  const int nwg = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 4096);
  k_copy_opt<double><<<nwg, WG_THREADS, 0, s>>>(n, v, st, out);
}

// ---------------------------------------------------------------------------
// Warm start from the previous steps: directions and their least-squares
// weights (the weights themselves are solved on the host from the Gram
// matrix of the initial residual and the directions' images under A).
// ---------------------------------------------------------------------------
// d = a - b, zero in the bMean cell
__global__ __launch_bounds__(WG_THREADS) void k_guess_direction(long long n,
                                                                const double* __restrict__ a,
                                                                const double* __restrict__ b,
                                                                long long row,
                                                                double* __restrict__ d) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    d[i] = i == row ? 0.0 : a[i] - b[i];
}

// Gram matrix of up to GUESS_MAXV vectors (upper triangle, row by row:
// (0,0), (0,1), ..., (1,1), ...), fixed-order partial sums; NV vectors.
template <int NV>
__global__ __launch_bounds__(WG_THREADS) void k_guess_gram(long long n, GuessVecs r,
                                                           double* part) {
  constexpr int ND = NV * (NV + 1) / 2;
  double v[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) v[k] = 0.0;
  const long long base = (long long)blockIdx.x * (4 * BLEN);
  for (int e = threadIdx.x; e < 4 * BLEN; e += WG_THREADS) {
    const long long i = base + e;
    if (i < n) {
      double q[NV];
#pragma unroll
      for (int a = 0; a < NV; ++a) q[a] = r.v[a][i];
      int k = 0;
#pragma unroll
      for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int b = a; b < NV; ++b) v[k++] += q[a] * q[b];
    }
  }
  grid_store_partial<ND>(v, part);
}

// Sums the partials and scatters them into the GUESS_MAXV-wide triangle
// (the entries of absent vectors 0).
template <int NV>
__global__ __launch_bounds__(WG_THREADS) void k_guess_gram_fin(const double* part, int nparts,
                                                               double* dots) {
  constexpr int ND = NV * (NV + 1) / 2;
  double tot[ND];
  grid_sum_partials<ND>(part, nparts, tot);
  if (threadIdx.x == 0) {
    int k = 0, o = 0;
    for (int a = 0; a < GUESS_MAXV; ++a)
      for (int b = a; b < GUESS_MAXV; ++b, ++o) dots[o] = a < NV && b < NV ? tot[k++] : 0.0;
  }
}

struct GuessWeights {
  double w[HIST_MAX];
};

// x = x0 + sum_{k < m} g_k (h_{k-1} - h_k) + gnet (pred - x0), h_{-1} = x0;
// the bMean cell keeps x0's value (every direction is zero there)
__global__ __launch_bounds__(WG_THREADS) void k_guess_blend(long long n,
                                                            const double* __restrict__ x0,
                                                            GuessVecs h, GuessWeights g, int m,
                                                            const double* __restrict__ pred,
                                                            double gnet, long long row,
                                                            double* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS) {
    const double b = x0[i];
    double v = b, prev = b;
    for (int k = 0; k < m; ++k) {
      const double hk = h.v[k][i];
      v = fma(g.w[k], prev - hk, v);
      prev = hk;
    }
    if (gnet != 0.0) v = fma(gnet, pred[i] - b, v);
    x[i] = i == row ? b : v;
  }
}

// Streaming copy for the bandwidth ceiling report (nontemporal float4 loads
// and stores, grid-stride over a fixed grid).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(WG_THREADS) void k_stream_copy(long long n4,
                                                            const f32x4* __restrict__ a,
                                                            f32x4* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n4;
       i += (long long)gridDim.x * WG_THREADS)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

void launch_stream_copy(hipStream_t s, long long n4, const void* a, void* b) {
  k_stream_copy<<<1024, WG_THREADS, 0, s>>>(n4, static_cast<const f32x4*>(a),
                                            static_cast<f32x4*>(b));
}

// Caller-side projection after a solve (PressureProjection.cpp:404-431):
// x -= sum_b h_b^3 sum x / sum_b 512 h_b^3, the volume-weighted mean, with the
// sum in a fixed order.
__global__ __launch_bounds__(WG_THREADS) void k_wsum(long long n, const double* __restrict__ x,
                                                     const double* __restrict__ h3,
                                                     double* part) {
  double v[1] = {0.0};
  const long long i0 = (long long)blockIdx.x * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    double xv[8];
    load8(x + i0, xv);
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) s += xv[e];
    v[0] = h3[i0 / BLEN] * s;
  }
  grid_store_partial<1>(v, part);
}

__global__ __launch_bounds__(WG_THREADS) void k_wsum_fin(const double* part, int nparts,
                                                         double* out) {
  double tot[1];
  grid_sum_partials<1>(part, nparts, tot);
  if (threadIdx.x == 0) out[0] = tot[0];
}

__global__ __launch_bounds__(WG_THREADS) void k_sub_mean(long long n, double* __restrict__ x,
                                                         const double* __restrict__ wsum,
                                                         double inv_vol) {
  const double m = wsum[0] * inv_vol;
  for (long long i = (long long)blockIdx.x * WG_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * WG_THREADS)
    x[i] -= m;
}

void launch_remove_mean(hipStream_t s, long long n, double* x, const double* h3, double inv_vol,
                        double* part, double* wsum) {
  const int nwg = (int)((n + 4 * BLEN - 1) / (4 * BLEN));
  k_wsum<<<nwg, WG_THREADS, 0, s>>>(n, x, h3, part);
  k_wsum_fin<<<1, WG_THREADS, 0, s>>>(part, nwg, wsum);
  const int nsub = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 8192);
  k_sub_mean<<<nsub, WG_THREADS, 0, s>>>(n, x, wsum, inv_vol);
}


__global__ void k_zero_cell(double* __restrict__ x, long long row) { x[row] = 0.0; }

void launch_zero_cell(hipStream_t s, double* x, long long row) {
  if (row >= 0) k_zero_cell<<<1, 1, 0, s>>>(x, row);
}

void launch_guess_direction(hipStream_t s, long long n, const double* a, const double* b,
                            long long row, double* d) {
  const int nwg = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 8192);
  k_guess_direction<<<nwg, WG_THREADS, 0, s>>>(n, a, b, row, d);
}

template <int NV>
static void guess_gram_nv(hipStream_t s, long long n, const GuessVecs& r, double* part,
                          double* dots) {
  const int nwg = (int)((n + 4 * BLEN - 1) / (4 * BLEN));
  k_guess_gram<NV><<<nwg, WG_THREADS, 0, s>>>(n, r, part);
  k_guess_gram_fin<NV><<<1, WG_THREADS, 0, s>>>(part, nwg, dots);
}

void launch_guess_gram(hipStream_t s, long long n, const GuessVecs& r, int nv, double* part,
                       double* dots) {
  static_assert(GUESS_MAXV == 10, "one instance per vector count");
  switch (nv) {
    case 1: guess_gram_nv<1>(s, n, r, part, dots); break;
    case 2: guess_gram_nv<2>(s, n, r, part, dots); break;
    case 3: guess_gram_nv<3>(s, n, r, part, dots); break;
    case 4: guess_gram_nv<4>(s, n, r, part, dots); break;
    case 5: guess_gram_nv<5>(s, n, r, part, dots); break;
    case 6: guess_gram_nv<6>(s, n, r, part, dots); break;
    case 7: guess_gram_nv<7>(s, n, r, part, dots); break;
    case 8: guess_gram_nv<8>(s, n, r, part, dots); break;
    case 9: guess_gram_nv<9>(s, n, r, part, dots); break;
    default: guess_gram_nv<10>(s, n, r, part, dots); break;
  }
}

void launch_guess_blend(hipStream_t s, long long n, const double* x0, const GuessVecs& h,
                        const double* gt, int m, const double* pred, double gnet, long long row,
                        double* x) {
  GuessWeights g{};
  for (int k = 0; k < m && k < HIST_MAX; ++k) g.w[k] = gt[k];
  const int nwg = (int)std::min<long long>((n + WG_THREADS - 1) / WG_THREADS, 8192);
  k_guess_blend<<<nwg, WG_THREADS, 0, s>>>(n, x0, h, g, m < HIST_MAX ? m : HIST_MAX, pred, gnet,
                                           row, x);
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
