// Precision-generic kernels of the classical BiCGSTAB iteration, shared by
// the FP64 drop-in (kernels.hip) and the FP32 inner solver of the
// mixed-precision form (mixed.hip). Device code; include from .hip files only.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace cup3d {

// ---------------------------------------------------------------------------
// Scalar state machine (single thread of a last workgroup).
// ---------------------------------------------------------------------------

// Start of iteration k: the breakdown test, beta and the restart of
// BiCGSTAB.cu:477-515, and the iteration cap of :474-475.
__device__ inline void begin_iteration(SolverState* st, double rho, double rr) {
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

// Finalize kernels (one workgroup) work on a copy of the state in LDS: the
// copy is loaded by the first threads together with the partial sums (one
// memory round trip for both), thread 0 runs the scalar recurrences on it,
// and the copy is written back in parallel.
constexpr int STATE_WORDS = sizeof(SolverState) / 4;
static_assert(sizeof(SolverState) % 4 == 0 && STATE_WORDS <= WG_THREADS, "state copy");

template <int K>
__device__ __forceinline__ void fin_load(const SolverState* st, SolverState& S,
                                         const double* part, int nparts, double (&tot)[K]) {
  if (threadIdx.x < STATE_WORDS)
    reinterpret_cast<unsigned*>(&S)[threadIdx.x] =
        reinterpret_cast<const unsigned*>(st)[threadIdx.x];
  grid_sum_partials<K>(part, nparts, tot);  // ends with a barrier
}

__device__ __forceinline__ void fin_store(SolverState* st, const SolverState& S) {
  __syncthreads();
  if (threadIdx.x < STATE_WORDS)
    reinterpret_cast<unsigned*>(st)[threadIdx.x] =
        reinterpret_cast<const unsigned*>(&S)[threadIdx.x];
}

// Contiguous 8-element vector access (16-byte loads/stores).
template <typename T>
__device__ __forceinline__ void load8(const T* __restrict__ p, T (&v)[8]) {
  typedef T V4 __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / sizeof(T);
#pragma unroll
  for (int k = 0; k < 8 / W; ++k) {
    const V4 q = reinterpret_cast<const V4*>(p)[k];
#pragma unroll
    for (int e = 0; e < W; ++e) v[k * W + e] = q[e];
  }
}
template <typename T>
__device__ __forceinline__ void store8(T* __restrict__ p, const T (&v)[8]) {
  typedef T V4 __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int W = 16 / sizeof(T);
#pragma unroll
  for (int k = 0; k < 8 / W; ++k) {
    V4 q;
#pragma unroll
    for (int e = 0; e < W; ++e) q[e] = v[k * W + e];
    reinterpret_cast<V4*>(p)[k] = q;
  }
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
  const long long i0 = (long long)blockIdx.x * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    double bv[8], xv[8];
    load8(rhs + i0, bv);
    if (x0) load8(x0 + i0, xv);
    T bo[8], xo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if (!x0) xv[e] = 0.0;
      if (i0 + e == zero_row) bv[e] = xv[e] = 0.0;
      bo[e] = (T)bv[e];
      xo[e] = (T)xv[e];
      v[0] += bv[e] * bv[e];
    }
    store8(b + i0, bo);
    store8(x + i0, xo);
  }
  grid_store_partial<1>(v, part);
}

[[maybe_unused]] static __global__ __launch_bounds__(WG_THREADS) void k_getvec_fin(
    SolverState* st, const double* part, int nparts, double tol_b) {
  __shared__ SolverState S;
  double tot[1];
  fin_load<1>(st, S, part, nparts, tot);
  if (threadIdx.x == 0) {
    S.bnorm2 = tot[0];
    const double tb = tol_b * sqrt(tot[0]);
    if (tb > S.tol_abs) S.tol_abs = tb;
  }
  fin_store(st, S);
}

// ---------------------------------------------------------------------------
// Coarse-fine faces, matrix-free (PoissonSolverExp.cpp:167-217 with the
// interpolation of PoissonSolverExp.h:384-478).
//
// Ghost of a fine cell next to a coarser block: 2/3 p_close - 1/5 p_far +
// 8/15 T, T the Taylor value of the coarse face-layer cell cc with
//   T = sum W[a][b] P(cc + (a-2) e_t1 + (b-2) e_t2),
//   W = dd + s1 D1(k1) d + s2 d D1(k2) + D2(k1) d + d D2(k2) + s1 s2 M(k1) M(k2),
// k the stencil class of the tangential coarse index (one-sided next to a
// half-block edge: BWD at 3 and 7, FWD at 0 and 4), s = +-1 the side of the
// fine cell inside the coarse cell. D1/D2/M over three points (M: two):
//   D1 CEN (-1,0,1): -1/8 0 1/8   BWD (-2,-1,0): 1/8 -1/2 3/8   FWD (0,1,2): -3/8 1/2 -1/8
//   D2 all classes: 1/32 -1/16 1/32
//   M  CEN (-1,+1): -1/8 1/8      BWD (-1,0): -1/4 1/4           FWD (0,1): -1/4 1/4
// The fine side's row is h (ghost - p); the coarse side's row is the sum over
// its four fine cells of h_f (p_f,close - ghost_f). Both rows sum to zero, so
// FP32 uses differences to the row's own cell value (every weight sum of a
// Taylor value is 1).
// ---------------------------------------------------------------------------
enum { TK_CEN = 0, TK_BWD = 1, TK_FWD = 2 };

__device__ __forceinline__ int taylor_kind(int t) {
  return (t == BS - 1 || t == BS / 2 - 1) ? TK_BWD : ((t == 0 || t == BS / 2) ? TK_FWD : TK_CEN);
}

// sum W[a][b] (P(a-2, b-2) - ref); get(o1, o2) reads the face layer at
// tangential offsets from the centre cell
template <typename T, typename F>
__device__ __forceinline__ T taylor_sum(F&& get, int k1, int k2, T s1, T s2, T ref) {
  const int o1 = k1 == TK_CEN ? -1 : (k1 == TK_BWD ? -2 : 0);
  const int o2 = k2 == TK_CEN ? -1 : (k2 == TK_BWD ? -2 : 0);
  const T d1a[3] = {k1 == TK_CEN ? (T)(-0.125) : (k1 == TK_BWD ? (T)0.125 : (T)(-0.375)),
                    k1 == TK_CEN ? (T)0 : (k1 == TK_BWD ? (T)(-0.5) : (T)0.5),
                    k1 == TK_CEN ? (T)0.125 : (k1 == TK_BWD ? (T)0.375 : (T)(-0.125))};
  const T d1b[3] = {k2 == TK_CEN ? (T)(-0.125) : (k2 == TK_BWD ? (T)0.125 : (T)(-0.375)),
                    k2 == TK_CEN ? (T)0 : (k2 == TK_BWD ? (T)(-0.5) : (T)0.5),
                    k2 == TK_CEN ? (T)0.125 : (k2 == TK_BWD ? (T)0.375 : (T)(-0.125))};
  const T d2[3] = {(T)0.03125, (T)(-0.0625), (T)0.03125};
  T sum = get(0, 0) - ref;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sum += (s1 * d1a[i] + d2[i]) * (get(o1 + i, 0) - ref);
    sum += (s2 * d1b[i] + d2[i]) * (get(0, o2 + i) - ref);
  }
  const int m1a = k1 == TK_FWD ? 0 : -1, m1b = k1 == TK_BWD ? 0 : 1;
  const int m2a = k2 == TK_FWD ? 0 : -1, m2b = k2 == TK_BWD ? 0 : 1;
  const T w1 = k1 == TK_CEN ? (T)0.125 : (T)0.25, w2 = k2 == TK_CEN ? (T)0.125 : (T)0.25;
  const T sm = s1 * s2 * w1 * w2;  // M(k) = w (-1 at ma, +1 at mb)
  sum += sm * (((get(m1b, m2b) - ref) - (get(m1a, m2b) - ref)) -
               ((get(m1b, m2a) - ref) - (get(m1a, m2a) - ref)));
  return sum;
}

// Contribution of coarse-fine face f of block b to the rows of its 64 face
// cells (lane = cell (t1, t2) = (lane & 7, lane >> 3)), added into ys.
// xs holds block b, pl is this wavefront's 64-value staging plane; nb0 is the
// face's first partner (the coarse block), nbq the finer block of this
// lane's quadrant (both prefetched by the caller).
__device__ __forceinline__ size_t coarse_plane_index(int f, int cb, int lane) {
  // cell (lane & 7, lane >> 3) of the face layer of coarse block cb that
  // touches face f of the finer block
  const int d = f >> 1;
  const bool minus = (f & 1) == 0;
  const int st1 = d == 0 ? 8 : (d == 1 ? 64 : 1), st2 = d == 0 ? 64 : (d == 1 ? 1 : 8);
  const int sd = d == 0 ? 1 : (d == 1 ? 8 : 64);
  return (size_t)cb * BLEN + (minus ? 7 : 0) * sd + (lane & 7) * st1 + (lane >> 3) * st2;
}

template <typename T>
__device__ __forceinline__ void iface_face(const DevGrid<T>& g, const T* __restrict__ x,
                                           const T* xs, T* pl, T* ys, int b, int f, int fk,
                                           int par, int lane, int nb0, int nbq, bool have_plane,
                                           T plane) {
  const int d = f >> 1;
  const bool minus = (f & 1) == 0;
  const int t1 = d == 0 ? 1 : (d == 1 ? 2 : 0), t2 = d == 0 ? 2 : (d == 1 ? 0 : 1);
  const int st1 = t1 == 0 ? 1 : (t1 == 1 ? 8 : 64), st2 = t2 == 0 ? 1 : (t2 == 1 ? 8 : 64);
  const int sd = d == 0 ? 1 : (d == 1 ? 8 : 64);
  const int u = lane & 7, v = lane >> 3;
  const int cl = (minus ? 0 : 7) * sd + u * st1 + v * st2;  // this lane's face cell
  const T xc = xs[cl];
  const bool diff = sizeof(T) == 4;
  const T ref = diff ? xc : (T)0;
  T acc = 0;
  if (fk == FACE_COARSE) {
    // stage the coarse block's face layer, [t2][t1] (prefetched by the
    // caller for the first two such faces)
    pl[lane] = have_plane ? plane : x[coarse_plane_index(f, nb0, lane)];
    wave_sync();
    const T h = g.h[b];
    const int c1 = (u >> 1) + (((par >> t1) & 1) ? BS / 2 : 0);
    const int c2 = (v >> 1) + (((par >> t2) & 1) ? BS / 2 : 0);
    const T s1 = (u & 1) ? (T)1 : (T)(-1), s2 = (v & 1) ? (T)1 : (T)(-1);
    const T tay = taylor_sum<T>([&](int o1, int o2) { return pl[(c2 + o2) * 8 + c1 + o1]; },
                                taylor_kind(c1), taylor_kind(c2), s1, s2, ref);
    const T xin = xs[cl + (minus ? sd : -sd)];
    if (diff)
      acc = h * ((T)(8.0 / 15.0) * tay - (xin - xc) * (T)0.2);
    else
      acc = h * ((T)(8.0 / 15.0) * tay - xin * (T)0.2 - xc * (T)(1.0 / 3.0));
  } else {
    // four fine cells of the finer block that covers this cell's quadrant
    const int fb = nbq;
    const T hf = g.h[fb];
    const size_t fo = (size_t)fb * BLEN;
    const int lc = minus ? 7 : 0, lf = minus ? 6 : 1;
    T xfc[4], xff[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f1 = (u & 3) * 2 + (q & 1), f2 = (v & 3) * 2 + (q >> 1);
      xfc[q] = x[fo + lc * sd + f1 * st1 + f2 * st2];
      xff[q] = x[fo + lf * sd + f1 * st1 + f2 * st2];
    }
    // The four fine ghosts share this cell's Taylor expansion and differ in
    // the signs s1, s2 = +-1 of the first-derivative and mixed terms, which
    // therefore cancel in the sum over the four: sum_q T_q = 4 (P + D2 terms).
    const int k1 = taylor_kind(u), k2 = taylor_kind(v);
    const int o1 = k1 == TK_CEN ? -1 : (k1 == TK_BWD ? -2 : 0);
    const int o2 = k2 == TK_CEN ? -1 : (k2 == TK_BWD ? -2 : 0);
    auto own = [&](int a1, int a2) { return xs[cl + a1 * st1 + a2 * st2] - ref; };
    const T tsum = (T)4 * (own(0, 0) + (T)0.03125 * ((own(o1, 0) + own(o1 + 2, 0)) +
                                                    (own(0, o2) + own(0, o2 + 2))) -
                           (T)0.0625 * (own(o1 + 1, 0) + own(0, o2 + 1)));
    T sc = 0, sf = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sc += xfc[q] - ref;
      sf += xff[q] - ref;
    }
    acc = hf * (sc * (T)(1.0 / 3.0) + sf * (T)0.2 - (T)(8.0 / 15.0) * tsum);
  }
  ys[cl] += acc;
  wave_sync();
}

// Loads and stores of vectors a kernel touches once: nontemporal when the
// vectors are too large to stay in the Infinity Cache between kernels.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Vectors of at least this many bytes are streamed with nontemporal access
// by the hot SpMV (K2, K4) and preconditioner (K1, K3) kernels. Measured on
// MI355X (256 MB Infinity Cache): with it, K2 on the 30.2 M-cell grid (FP32
// vectors of 121 MB) 97 -> 90 us and K1 146 -> 134 us; on a 4.19 M-cell
// grid (17 MB FP32, 34 MB FP64 vectors) it made K2 12-15 % slower, the
// vectors no longer found in the cache by the next kernel.
constexpr size_t STREAM_MIN_BYTES = size_t(64) << 20;

// ---------------------------------------------------------------------------
// Matrix-free operator y = A x (PoissonSolverExp.cpp:305-367 + interface rows)
// with the bMean handling of hd_cusparseSpMV (BiCGSTAB.cu:392-413) and a
// fused reduction epilogue.
//
// One wavefront per block: lane = (ix, iy), each lane walks its z column in
// registers; the block is staged in LDS for the x/y neighbours. Every global
// load (the column, the epilogue operand, the six face halos of same-level
// neighbours) is issued before the LDS staging, so a wave waits for memory
// once. Grids with coarse-fine faces (IFACE) stage y in LDS once more to add
// the interface rows. A wavefront reads only its own LDS slices, so it
// synchronises with itself alone (wave_sync) and never waits for the other
// three; the workgroup barrier is the partial-sum reduction at the end.
// ---------------------------------------------------------------------------
//
// HELM (FP64, bMean 0): the implicit-diffusion operator of one velocity
// component, KernelLHSDiffusion (AdvectionDiffusionImplicit.h:40-56): the
// same stencil and coarse-fine faces, plus -h^3 sigma x and the Dirichlet
// ghosts of BlockLabBC (Definitions.h:58-122,287-330) on the boundary faces
// selected by g.dmask, both folded into the diagonal of the cell.
template <typename T, int MODE, int BM, bool IFACE, bool HELM = false, bool NT = false>
__global__ __launch_bounds__(WG_THREADS) void k_spmv(DevGrid<T> g, const T* __restrict__ x,
                                                     T* __restrict__ y, const T* __restrict__ a,
                                                     T* __restrict__ c, T* __restrict__ c2,
                                                     SolverState* st, double* part) {
  constexpr bool NEED_A = MODE != SPMV_PLAIN;
  // sum_b h_b^3 sum x feeds the bMean = 1 row and the bMean = 2 term only
  constexpr bool NEED_SUM = BM == 1 || BM == 2;
  __shared__ T xs[4][BLEN];
  __shared__ T ys[IFACE ? 4 : 1][IFACE ? BLEN : 1];
  __shared__ T pl[IFACE ? 4 : 1][IFACE ? 64 : 1];
  // per wavefront: the face partners fnb[24 b .. 24 b + 23], the six face
  // kinds and the index parities of its block
  __shared__ int fmeta[IFACE ? 4 : 1][IFACE ? 32 : 1];
  // with coarse-fine faces the epilogue operand waits in LDS, not registers
  constexpr bool A_LDS = IFACE && NEED_A;
  __shared__ T as[A_LDS ? 4 : 1][A_LDS ? BLEN : 1];
  if (MODE != SPMV_PLAIN && MODE != SPMV_RESID && st->status != ST_RUNNING) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ix = lane & 7, iy = lane >> 3;
  const int b = xcd_remap(blockIdx.x, gridDim.x) * 4 + w;
  const bool valid = b < g.nb;
  const size_t off = (size_t)b * BLEN;
  T col[BS], av[BS], hx[BS], hy[BS], zlo = 0, zhi = 0;
  T h = 0, pre0 = 0, pre1 = 0;
  int meta = 0, cf0 = -1, cf1 = -1;
  if (valid) {
    // face partners, kinds and parities: one word per lane of the block's
    // record, loaded first and by every lane (lanes 32-63 repeat 0-31), so
    // that the load is not predicated and its value arrives first
    if (IFACE) meta = g.bmeta[(size_t)b * 32 + (lane & 31)];
    const int* nb6 = g.nbr + b * 6;
    const int nxm = nb6[0], nxp = nb6[1], nym = nb6[2], nyp = nb6[3], nzm = nb6[4], nzp = nb6[5];
    h = g.h[b];
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) col[iz] = x[off + iz * 64 + lane];
    if (NEED_A) {
#pragma unroll
      for (int iz = 0; iz < BS; ++iz) av[iz] = ld_stream<NT>(&a[off + iz * 64 + lane]);
    }
    // face halos: a lane is on at most one x face and one y face. A lane on
    // a face with a same-level neighbour reads the neighbour's face cell;
    // every other lane reads its own cell (the zero-flux ghost of a domain
    // face, unused inside the block). The addresses are selected, not the
    // loads, so every halo load issues with the column without waiting for it.
    const int nx = ix == 0 ? nxm : (ix == 7 ? nxp : -1);
    const int sx = ix == 0 ? 7 : 0;
    const int ny = iy == 0 ? nym : (iy == 7 ? nyp : -1);
    const int sy = iy == 0 ? 7 : 0;
    const T* px = x + (nx >= 0 ? (size_t)nx * BLEN + iy * 8 + sx : off + lane);
    const T* py = x + (ny >= 0 ? (size_t)ny * BLEN + sy * 8 + ix : off + lane);
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      hx[iz] = px[iz * 64];
      hy[iz] = py[iz * 64];
    }
    zlo = x[nzm >= 0 ? (size_t)nzm * BLEN + 7 * 64 + lane : off + lane];
    zhi = x[nzp >= 0 ? (size_t)nzp * BLEN + lane : off + 7 * 64 + lane];
    if (IFACE) {
      // the face layers of the coarser neighbours of the first two faces
      // that have one, loaded now so that their latency overlaps the stencil
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (__builtin_amdgcn_readlane(meta, 24 + f) == FACE_COARSE) {
          if (cf0 < 0) cf0 = f;
          else if (cf1 < 0) cf1 = f;
        }
      }
      if (cf0 >= 0) pre0 = x[coarse_plane_index(cf0, __builtin_amdgcn_readlane(meta, 4 * cf0), lane)];
      if (cf1 >= 0) pre1 = x[coarse_plane_index(cf1, __builtin_amdgcn_readlane(meta, 4 * cf1), lane)];
    }
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) xs[w][iz * 64 + lane] = col[iz];
    if (A_LDS) {
#pragma unroll
      for (int iz = 0; iz < BS; ++iz) as[w][iz * 64 + lane] = av[iz];
    }
    if (IFACE) fmeta[w][lane & 31] = meta;
  }
  wave_sync();
  T yv[BS];
  double msum = 0.0;
  if (valid) {
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      const T c0 = col[iz];
      const int li = iz * 64 + lane;
      // every lane reads LDS (a face lane its own cell) and then selects, so
      // the reads are not predicated one by one
      const T lxm = xs[w][li - (ix > 0 ? 1 : 0)], lxp = xs[w][li + (ix < 7 ? 1 : 0)];
      const T lym = xs[w][li - (iy > 0 ? 8 : 0)], lyp = xs[w][li + (iy < 7 ? 8 : 0)];
      const T xm = ix > 0 ? lxm : hx[iz];
      const T xp = ix < 7 ? lxp : hx[iz];
      const T ym = iy > 0 ? lym : hy[iz];
      const T yp = iy < 7 ? lyp : hy[iz];
      const T zm = iz > 0 ? col[iz - 1] : zlo;
      const T zp = iz < BS - 1 ? col[iz + 1] : zhi;
      if (sizeof(T) == 8)
        yv[iz] = h * ((((xm + xp) + (ym + yp)) + (zm + zp)) - (T)6 * c0);
      else  // FP32: differences of neighbouring values are exact (Sterbenz), so
            // the result keeps full relative precision on smooth fields
        yv[iz] = h * ((((xm - c0) + (xp - c0)) + ((ym - c0) + (yp - c0))) +
                      ((zm - c0) + (zp - c0)));
      if (NEED_SUM) msum += (double)c0;
    }
    if (HELM) {
      // domain-boundary faces of this block with a Dirichlet ghost
      int dm = 0;
#pragma unroll
      for (int f = 0; f < 6; ++f) dm |= (g.fkind[6 * b + f] == FACE_NONE ? 1 : 0) << f;
      dm &= g.dmask;
      const T shift = -(h * h * h) * (T)g.helm_sigma;
      const int nxy = (ix == 0 ? (dm & 1) : 0) + (ix == BS - 1 ? ((dm >> 1) & 1) : 0) +
                      (iy == 0 ? ((dm >> 2) & 1) : 0) + (iy == BS - 1 ? ((dm >> 3) & 1) : 0);
#pragma unroll
      for (int iz = 0; iz < BS; ++iz) {
        const int nd = nxy + (iz == 0 ? ((dm >> 4) & 1) : 0) + (iz == BS - 1 ? ((dm >> 5) & 1) : 0);
        yv[iz] += (shift - (T)(2 * nd) * h) * col[iz];
      }
    }
  }
  if (IFACE) {
    if (valid) {
#pragma unroll
      for (int iz = 0; iz < BS; ++iz) ys[w][iz * 64 + lane] = yv[iz];
    }
    wave_sync();
    if (valid) {
      const int par = fmeta[w][30];
      const int fq = (ix >= BS / 2 ? 1 : 0) + (iy >= BS / 2 ? 2 : 0);
      // one face at a time: the face code is instantiated once, which keeps
      // the register count (and so the occupancy) close to the uniform kernel's
#pragma unroll 1
      for (int f = 0; f < 6; ++f) {
        const int fk = __builtin_amdgcn_readfirstlane(fmeta[w][24 + f]);
        if (fk >= FACE_COARSE)
          iface_face<T>(g, x, xs[w], pl[w], ys[w], b, f, fk, par, lane, fmeta[w][4 * f],
                        fmeta[w][4 * f + fq], f == cf0 || f == cf1, f == cf0 ? pre0 : pre1);
      }
    }
    wave_sync();
    if (valid) {
#pragma unroll
      for (int iz = 0; iz < BS; ++iz) yv[iz] = ys[w][iz * 64 + lane];
    }
  }
  const long long row = g.bmean_row;  // origin cell for BM 1 and 3, else -1
  const bool origin = valid && row == (long long)off;
  if (BM == 3 && origin && lane == 0) yv[0] = h * col[0];  // row h*I (.cpp:311-318)
  // epilogue: stores and partial sums
  double v[3] = {0.0, 0.0, 0.0};
  if (valid) {
    if (NEED_SUM) v[0] = (double)h * (double)h * (double)h * msum;
#pragma unroll
    for (int iz = 0; iz < BS; ++iz) {
      const size_t i = off + iz * 64 + lane;
      // the bMean row is written by the last workgroup only, so no two
      // workgroups (possibly on different XCDs) ever store the same address
      if ((BM == 1) && origin && lane == 0 && iz == 0) continue;
      const T yy = yv[iz];
      if (A_LDS) av[iz] = as[w][iz * 64 + lane];
      if (MODE == SPMV_PLAIN) {
        st_stream<NT>(&y[i], yy);
      } else if (MODE == SPMV_INIT) {
        st_stream<NT>(&y[i], yy);
        const T rr = av[iz] - yy;
        c[i] = rr;
        c2[i] = rr;
        v[1] += (double)yy * (double)yy;
        v[2] += (double)rr * (double)rr;
      } else if (MODE == SPMV_DOT1) {
        st_stream<NT>(&y[i], yy);
        v[1] += (double)av[iz] * (double)yy;
      } else if (MODE == SPMV_DOT2) {
        st_stream<NT>(&y[i], yy);
        v[1] += (double)av[iz] * (double)yy;
        v[2] += (double)yy * (double)yy;
      } else if (MODE == SPMV_RESID) {
        const T rr = av[iz] - yy;
        c[i] = rr;
        v[1] += (double)rr * (double)rr;
      }
    }
  }
  grid_store_partial<3>(v, part);
}

// Finalize of k_spmv (one workgroup): totals, the bMean = 1 row
// (y[row] = sum_b h_b^3 sum x, BiCGSTAB.cu:406-407) and the scalar updates.
template <typename T, int MODE, int BM>
__global__ __launch_bounds__(WG_THREADS) void k_spmv_fin(long long row, T* __restrict__ y,
                                                         const T* __restrict__ a,
                                                         T* __restrict__ c, T* __restrict__ c2,
                                                         SolverState* st, const double* part,
                                                         int nparts) {
  __shared__ SolverState S;
  double tot[3];
  fin_load<3>(st, S, part, nparts, tot);
  if (MODE != SPMV_PLAIN && MODE != SPMV_RESID && S.status != ST_RUNNING) return;
  if (threadIdx.x == 0) {
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
    S.mean_sum = tot[0];
    if (MODE == SPMV_INIT) {
      // initial residual (BiCGSTAB.cu:435-471)
      S.nu2 = tot[1];
      S.err = sqrt(tot[2]);
      S.err_init = S.err;
      S.err_opt = S.err;
      S.rhrh = tot[2];
      begin_iteration(&S, tot[2], tot[2]);
    } else if (MODE == SPMV_DOT1) {
      S.dot_a = tot[1];
      S.alpha = S.rho_curr / (tot[1] + S.eps);  // set_alpha (:276-279)
      S.p_zero = 0;
      S.restart_pending = 0;
    } else if (MODE == SPMV_DOT2) {
      S.dot_ts = tot[1];
      S.dot_tt = tot[2];
      S.omega = tot[1] / (tot[2] + S.eps);  // set_omega (:281-284)
    } else if (MODE == SPMV_RESID) {
      S.true_rr = tot[1];
    }
  }
  fin_store(st, S);
}

// bMeanConstraint = 2 (hd_cusparseSpMV, BiCGSTAB.cu:392-411):
// y = A x + (sum_b h_b^3 sum x) h^3 in every cell. The total is known only
// once k_spmv has run, so k_spmv stores the plain A x and its partial
// h^3-sums; k_bmean2_sum reduces them into st->mean_sum, and k_bmean2_apply
// adds the rank-one term and redoes the epilogue (outputs and partial sums)
// on the corrected y, for k_spmv_fin to finish as for bMean = 0.
template <int MODE>
__global__ __launch_bounds__(WG_THREADS) void k_bmean2_sum(SolverState* st, const double* part,
                                                           int nparts) {
  __shared__ SolverState S;
  double tot[3];
  fin_load<3>(st, S, part, nparts, tot);
  if (MODE != SPMV_PLAIN && MODE != SPMV_RESID && S.status != ST_RUNNING) return;
  if (threadIdx.x == 0) S.mean_sum = tot[0];
  fin_store(st, S);
}

template <typename T, int MODE>
__global__ __launch_bounds__(WG_THREADS) void k_bmean2_apply(DevGrid<T> g, T* __restrict__ y,
                                                             const T* __restrict__ a,
                                                             T* __restrict__ c,
                                                             T* __restrict__ c2,
                                                             SolverState* st, double* part) {
  if (MODE != SPMV_PLAIN && MODE != SPMV_RESID && st->status != ST_RUNNING) return;
  const double red = st->mean_sum;
  // one thread per 8 consecutive cells (never across a block)
  const long long j = (long long)blockIdx.x * WG_THREADS + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (j < (long long)g.nb * (BLEN / 8)) {
    const size_t i = (size_t)j * 8;
    const double hb = (double)g.h[i / BLEN];
    const T corr = (T)(red * (hb * hb * hb));
    if (MODE == SPMV_RESID) {  // only c = a - A x was stored
      T cv[8];
      load8(c + i, cv);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        cv[e] -= corr;
        v[1] += (double)cv[e] * (double)cv[e];
      }
      store8(c + i, cv);
    } else {
      T yv[8], av[8];
      load8(y + i, yv);
#pragma unroll
      for (int e = 0; e < 8; ++e) yv[e] += corr;
      store8(y + i, yv);
      if (MODE != SPMV_PLAIN) load8(a + i, av);
      if (MODE == SPMV_INIT) {
        T rr[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          rr[e] = av[e] - yv[e];
          v[1] += (double)yv[e] * (double)yv[e];
          v[2] += (double)rr[e] * (double)rr[e];
        }
        store8(c + i, rr);
        store8(c2 + i, rr);
      } else if (MODE == SPMV_DOT1 || MODE == SPMV_DOT2) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          v[1] += (double)av[e] * (double)yv[e];
          if (MODE == SPMV_DOT2) v[2] += (double)yv[e] * (double)yv[e];
        }
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) v[0] = red;  // k_spmv_fin keeps it as mean_sum
  grid_store_partial<3>(v, part);
}

// Launch for a grid: picks the bMean and interface variants.
// An event, if given, is recorded right after the main SpMV kernel.
template <typename T, int MODE, int BM, bool HELM, bool NT>
inline void spmv_kernel(hipStream_t s, int nwg, const DevGrid<T>& g, const T* x, T* y,
                        const T* a, T* c, T* c2, SolverState* st, double* part) {
  if (g.has_iface)
    k_spmv<T, MODE, BM, true, HELM, NT><<<nwg, WG_THREADS, 0, s>>>(g, x, y, a, c, c2, st, part);
  else
    k_spmv<T, MODE, BM, false, HELM, NT><<<nwg, WG_THREADS, 0, s>>>(g, x, y, a, c, c2, st, part);
}
template <typename T, int MODE, int BM, bool HELM = false>
inline void spmv_launch_iface(hipStream_t s, const DevGrid<T>& g, const T* x, T* y, const T* a,
                              T* c, T* c2, SolverState* st, double* part,
                              hipEvent_t mid = nullptr) {
  int nwg = (g.nb + 3) / 4;
  bool launched = false;
  if constexpr (MODE == SPMV_DOT1 || MODE == SPMV_DOT2) {  // the per-iteration SpMVs
    if ((size_t)g.n * sizeof(T) >= STREAM_MIN_BYTES) {
      spmv_kernel<T, MODE, BM, HELM, true>(s, nwg, g, x, y, a, c, c2, st, part);
      launched = true;
    }
  }
  if (!launched) spmv_kernel<T, MODE, BM, HELM, false>(s, nwg, g, x, y, a, c, c2, st, part);
  if (mid) (void)hipEventRecord(mid, s);
  if (BM == 2) {
    k_bmean2_sum<MODE><<<1, WG_THREADS, 0, s>>>(st, part, nwg);
    // nb * 64 threads of 8 cells: (nb + 3) / 4 workgroups again, so the
    // partial buffer sized for k_spmv holds them
    nwg = (int)(((long long)g.nb * (BLEN / 8) + WG_THREADS - 1) / WG_THREADS);
    k_bmean2_apply<T, MODE><<<nwg, WG_THREADS, 0, s>>>(g, y, a, c, c2, st, part);
  }
  k_spmv_fin<T, MODE, BM><<<1, WG_THREADS, 0, s>>>(g.bmean_row, y, a, c, c2, st, part, nwg);
}
template <typename T, int MODE>
inline void spmv_launch(hipStream_t s, const DevGrid<T>& g, const T* x, T* y, const T* a, T* c,
                        T* c2, SolverState* st, double* part, hipEvent_t mid = nullptr) {
  if constexpr (sizeof(T) == 8) {
    if (g.helm) {  // implicit diffusion: no bMean row
      spmv_launch_iface<T, MODE, 0, true>(s, g, x, y, a, c, c2, st, part, mid);
      return;
    }
  }
  if (g.bmean == 1) spmv_launch_iface<T, MODE, 1>(s, g, x, y, a, c, c2, st, part, mid);
  else if (g.bmean == 2) spmv_launch_iface<T, MODE, 2>(s, g, x, y, a, c, c2, st, part, mid);
  else if (g.bmean == 3) spmv_launch_iface<T, MODE, 3>(s, g, x, y, a, c, c2, st, part, mid);
  else spmv_launch_iface<T, MODE, 0>(s, g, x, y, a, c, c2, st, part, mid);
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

// A K5 workgroup covers CH runs of 2048 elements. On grids of at least 8192
// runs CH = 2 halves the partial sums the finalize kernel reads (30 M cells:
// 185 us against 191 us with it); smaller grids keep every CU busy with
// CH = 1 (4.2 M cells, FP64: 48.7 us against 54.6 us).
template <typename T, int CH>
__global__ __launch_bounds__(WG_THREADS) void k_update(long long n, KrylovVectors<T> v,
                                                       SolverState* st, double* part) {
  if (st->status != ST_RUNNING) return;
  const int xc = st->xcur, xo = st->xopt, xn = third_buffer(xc, xo);
  const T alpha = (T)st->alpha, omega = (T)st->omega;
  double s[2] = {0.0, 0.0};
  // each thread owns 8 consecutive elements of each of this workgroup's runs
#pragma unroll
  for (int c = 0; c < CH; ++c) {
  const long long i0 = ((long long)blockIdx.x * CH + c) * (4 * BLEN) + 8 * threadIdx.x;
  if (i0 < n) {
    T xv[8], z1[8], z2[8], t[8], r[8], rh[8];
    load8(v.x[xc] + i0, xv);
    load8(v.z1 + i0, z1);
    load8(v.z2 + i0, z2);
    load8(v.t + i0, t);
    load8(v.r + i0, r);
    load8(v.rhat + i0, rh);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      xv[e] = fma(omega, z2[e], fma(alpha, z1[e], xv[e]));
      r[e] = fma(-omega, t[e], r[e]);
      s[0] += (double)rh[e] * (double)r[e];
      s[1] += (double)r[e] * (double)r[e];
    }
    store8(v.x[xn] + i0, xv);
    store8(v.r + i0, r);
  }
  }
  grid_store_partial<2>(s, part);
}

// Finalize of k_update (one workgroup): |r|, x_opt, convergence test, restart
// and the next iteration's scalars (BiCGSTAB.cu:580-607).
// End of an iteration (BiCGSTAB.cu:584-607): the residual norm, x_opt,
// convergence, then the start of the next iteration.
__device__ inline void update_scalars(SolverState& S, double rho, double rr) {
  const int xn = third_buffer(S.xcur, S.xopt);
  const double err = sqrt(rr);
  S.err = err;
  S.iters_done = S.iter;
  S.xcur = xn;
  if (!(err == err) || err > 1.7e308) {
    S.status = ST_NONFINITE;
    return;
  }
  if (err < S.err_opt) {
    S.err_opt = err;
    S.xopt = xn;
    if (err <= S.tol_abs || err / S.err_init <= S.tol_rel) {
      S.status = ST_CONVERGED;
      return;
    }
  }
  S.rho_prev = S.rho_curr;  // set_rho (:607)
  begin_iteration(&S, rho, rr);
}

static __global__ __launch_bounds__(WG_THREADS) void k_update_fin(SolverState* st, const double* part,
                                                           int nparts) {
  __shared__ SolverState S;
  double tot[2];
  fin_load<2>(st, S, part, nparts, tot);
  if (S.status != ST_RUNNING) return;
  if (threadIdx.x == 0) update_scalars(S, tot[0], tot[1]);
  fin_store(st, S);
}

// K5 and its finalize kernel (an event, if given, is recorded between them).
template <typename T>
inline void launch_k5(hipStream_t s, long long n, const KrylovVectors<T>& v, SolverState* st,
                      double* part, hipEvent_t mid = nullptr) {
  const long long runs = (n + 4 * BLEN - 1) / (4 * BLEN);
  const int ch = runs >= 8192 ? 2 : 1;
  const int nwg = (int)((runs + ch - 1) / ch);
  if (ch == 2) k_update<T, 2><<<nwg, WG_THREADS, 0, s>>>(n, v, st, part);
  else k_update<T, 1><<<nwg, WG_THREADS, 0, s>>>(n, v, st, part);
  if (mid) (void)hipEventRecord(mid, s);
  k_update_fin<<<1, WG_THREADS, 0, s>>>(st, part, nwg);
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

}  // namespace cup3d
