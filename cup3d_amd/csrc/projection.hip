// The caller around the solve, on the device: the rhs of the pressure
// equation (KernelPressureRHS, source/operators/PressureProjection.cpp:96-199)
// and the pressure-gradient velocity update (KernelGradP :244-323 and
// :449-466), with the flux correction at coarse-fine faces, over Cubism's
// block layout (velocities [nb][8][8][8][3], scalars [nb][8][8][8]).
//
// Lab (ghost) values one cell outside a block, per face kind:
//   same level      the neighbour's cell (periodic wrap through the tables);
//   coarser         2/3 f_close - 1/5 f_far + 8/15 T(f_coarse), the fine ghost
//                   of the operator's interface interpolation (exact for
//                   quadratics; the Taylor value T as the SpMV evaluates it);
//   domain boundary velocity (BlockLabBC, Definitions.h:284-435): the normal
//                   component mirrored, ghost = -u, for walls and freespace;
//                   pressure (BlockLabNeumann3D): ghost = p;
//   finer           the coarse cell carries no face term of its own and gets
//                   the sum of the four fine cells' face terms instead, each
//                   with its own fine ghost (Cubism's flux correction).
// A face term is s fac (f_nb + f_c) for side s = -1 / +1, so the interior
// formula fac (f(x+h) - f(x-h)) is the sum of a cell's two face terms.
//
// One thread per cell: these kernels run once per time step, against several
// hundred solver iterations, and read each block's neighbours through L2.

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"
#include "krylov.h"

namespace cup3d {

namespace {

__device__ __forceinline__ long long cidx(int b, const int c[3]) {
  return (long long)b * BLEN + c[2] * 64 + c[1] * 8 + c[0];
}

__device__ __forceinline__ void face_tangents(int d, int& t1, int& t2) {
  t1 = d == 0 ? 1 : (d == 1 ? 2 : 0);
  t2 = d == 0 ? 2 : (d == 1 ? 0 : 1);
}

// Ghost of component `comp` of an `nc`-component field across face f of the
// fine cell c of block b, whose neighbour there is coarser.
__device__ double coarse_ghost(const DevGrid<double>& g, const double* __restrict__ F, int nc,
                               int comp, int b, const int c[3], int f) {
  const int d = f >> 1;
  const bool minus = (f & 1) == 0;
  int t1, t2;
  face_tangents(d, t1, t2);
  const int cb = g.fnb[(size_t)(6 * b + f) * 4];
  const int par = g.bpar[b];
  const int u = c[t1], v = c[t2];
  const int c1 = (u >> 1) + (((par >> t1) & 1) ? BS / 2 : 0);
  const int c2 = (v >> 1) + (((par >> t2) & 1) ? BS / 2 : 0);
  const double s1 = (u & 1) ? 1.0 : -1.0, s2 = (v & 1) ? 1.0 : -1.0;
  const int cn = minus ? BS - 1 : 0;
  auto get = [&](int o1, int o2) {
    int q[3];
    q[d] = cn;
    q[t1] = c1 + o1;
    q[t2] = c2 + o2;
    return F[cidx(cb, q) * nc + comp];
  };
  const double T = taylor_sum<double>(get, taylor_kind(c1), taylor_kind(c2), s1, s2, 0.0);
  int ci[3] = {c[0], c[1], c[2]};
  ci[d] += minus ? 1 : -1;
  const double close = F[cidx(b, c) * nc + comp], far = F[cidx(b, ci) * nc + comp];
  return (2.0 / 3.0) * close - 0.2 * far + (8.0 / 15.0) * T;
}

// Sum over the four fine cells behind the finer face f of coarse cell c
// (block b) of fac_f ((G_q + F_q) - chi_q (Gs_q + Fs_q)), G the fine ghosts
// (Fs and chi may be null).
__device__ double fine_faces(const DevGrid<double>& g, const double* __restrict__ F,
                             const double* __restrict__ Fs, const double* __restrict__ chi,
                             int nc, int comp, int b, const int c[3], int f, double fac_f) {
  const int d = f >> 1;
  const bool minus = (f & 1) == 0;
  int t1, t2;
  face_tangents(d, t1, t2);
  const int u = c[t1], v = c[t2];
  const int q = (u >= BS / 2 ? 1 : 0) + (v >= BS / 2 ? 2 : 0);
  const int fb = g.fnb[(size_t)(6 * b + f) * 4 + q];
  const int lc = minus ? BS - 1 : 0, lf = minus ? BS - 2 : 1;
  const int k1 = taylor_kind(u), k2 = taylor_kind(v);
  auto own = [&](const double* A) {
    return [=](int o1, int o2) {
      int p[3] = {c[0], c[1], c[2]};
      p[t1] += o1;
      p[t2] += o2;
      return A[cidx(b, p) * nc + comp];
    };
  };
  double acc = 0.0;
#pragma unroll 1
  for (int qq = 0; qq < 4; ++qq) {
    int fc[3], ff[3];
    fc[d] = lc;
    ff[d] = lf;
    fc[t1] = ff[t1] = (u & 3) * 2 + (qq & 1);
    fc[t2] = ff[t2] = (v & 3) * 2 + (qq >> 1);
    const double s1 = (qq & 1) ? 1.0 : -1.0, s2 = (qq >> 1) ? 1.0 : -1.0;
    const long long kc = cidx(fb, fc), kf = cidx(fb, ff);
    const double T = taylor_sum<double>(own(F), k1, k2, s1, s2, 0.0);
    const double G = (2.0 / 3.0) * F[kc * nc + comp] - 0.2 * F[kf * nc + comp] + (8.0 / 15.0) * T;
    double t = G + F[kc * nc + comp];
    if (Fs && chi) {
      const double Ts = taylor_sum<double>(own(Fs), k1, k2, s1, s2, 0.0);
      const double Gs =
          (2.0 / 3.0) * Fs[kc * nc + comp] - 0.2 * Fs[kf * nc + comp] + (8.0 / 15.0) * Ts;
      t -= chi[kc] * (Gs + Fs[kc * nc + comp]);
    }
    acc += fac_f * t;
  }
  return acc;
}

// KernelPressureRHS: b = fac (uE - uW + vN - vS + wB - wF) - chi fac (the
// same of udef) + the finer faces' flux corrections, fac = h^2 / (2 dt).
__global__ __launch_bounds__(WG_THREADS) void k_pressure_rhs(DevGrid<double> g,
                                                             const double* __restrict__ vel,
                                                             const double* __restrict__ udef,
                                                             const double* __restrict__ chi,
                                                             double dt,
                                                             double* __restrict__ out) {
  const long long gid = (long long)blockIdx.x * WG_THREADS + threadIdx.x;
  if (gid >= g.n) return;
  const int b = (int)(gid / BLEN), cl = (int)(gid % BLEN);
  const int c[3] = {cl & 7, (cl >> 3) & 7, cl >> 6};
  const double h = g.h[b], fac = 0.5 * h * h / dt;
  const double hf = 0.5 * h, fac_f = 0.5 * hf * hf / dt;
  const bool obst = udef && chi;
  double U[6], Us[6], corr = 0.0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f >> 1, s = (f & 1) ? 1 : -1;
    int cn[3] = {c[0], c[1], c[2]};
    cn[d] += s;
    Us[f] = 0.0;
    if (cn[d] >= 0 && cn[d] < BS) {
      U[f] = vel[cidx(b, cn) * 3 + d];
      if (obst) Us[f] = udef[cidx(b, cn) * 3 + d];
      continue;
    }
    const int fk = g.fkind[6 * b + f];
    if (fk == FACE_SAME) {
      cn[d] = s < 0 ? BS - 1 : 0;
      const int nb = g.nbr[6 * b + f];
      U[f] = vel[cidx(nb, cn) * 3 + d];
      if (obst) Us[f] = udef[cidx(nb, cn) * 3 + d];
    } else if (fk == FACE_COARSE) {
      U[f] = coarse_ghost(g, vel, 3, d, b, c, f);
      if (obst) Us[f] = coarse_ghost(g, udef, 3, d, b, c, f);
    } else {  // domain boundary or finer neighbours: no face term of its own
      U[f] = -vel[gid * 3 + d];
      if (obst) Us[f] = -udef[gid * 3 + d];
      if (fk == FACE_FINE)
        corr += s * fine_faces(g, vel, obst ? udef : nullptr, obst ? chi : nullptr, 3, d, b, c,
                               f, fac_f);
    }
  }
  double v = fac * (U[1] - U[0] + U[3] - U[2] + U[5] - U[4]);
  if (obst) v += -chi[gid] * fac * (Us[1] - Us[0] + Us[3] - Us[2] + Us[5] - Us[4]);
  out[gid] = v + corr;
}

// KernelGradP and the velocity update: o = -dt h^2 / 2 (p(x+h) - p(x-h)) per
// component plus the finer faces' flux corrections; vel += o / h^3.
__global__ __launch_bounds__(WG_THREADS) void k_grad_p_update(DevGrid<double> g,
                                                              const double* __restrict__ p,
                                                              double dt,
                                                              double* __restrict__ vel) {
  const long long gid = (long long)blockIdx.x * WG_THREADS + threadIdx.x;
  if (gid >= g.n) return;
  const int b = (int)(gid / BLEN), cl = (int)(gid % BLEN);
  const int c[3] = {cl & 7, (cl >> 3) & 7, cl >> 6};
  const double h = g.h[b], fac = -0.5 * dt * h * h;
  const double hf = 0.5 * h, fac_f = -0.5 * dt * hf * hf;
  const double inv = 1.0 / (h * h * h);
  double P[6], corr[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int d = f >> 1, s = (f & 1) ? 1 : -1;
    int cn[3] = {c[0], c[1], c[2]};
    cn[d] += s;
    if (cn[d] >= 0 && cn[d] < BS) {
      P[f] = p[cidx(b, cn)];
      continue;
    }
    const int fk = g.fkind[6 * b + f];
    if (fk == FACE_SAME) {
      cn[d] = s < 0 ? BS - 1 : 0;
      P[f] = p[cidx(g.nbr[6 * b + f], cn)];
    } else if (fk == FACE_COARSE) {
      P[f] = coarse_ghost(g, p, 1, 0, b, c, f);
    } else if (fk == FACE_FINE) {
      P[f] = -p[gid];
      corr[d] += s * fine_faces(g, p, nullptr, nullptr, 1, 0, b, c, f, fac_f);
    } else {  // domain boundary: zero gradient
      P[f] = p[gid];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double o = fac * (P[2 * d + 1] - P[2 * d]) + corr[d];
    vel[gid * 3 + d] += inv * o;
  }
}

}  // namespace

void launch_pressure_rhs(hipStream_t s, const DevGrid<double>& g, const double* vel,
                         const double* udef, const double* chi, double dt, double* out) {
  const int nwg = (int)((g.n + WG_THREADS - 1) / WG_THREADS);
  k_pressure_rhs<<<nwg, WG_THREADS, 0, s>>>(g, vel, udef, chi, dt, out);
}

void launch_grad_p_update(hipStream_t s, const DevGrid<double>& g, const double* p, double dt,
                          double* vel) {
  const int nwg = (int)((g.n + WG_THREADS - 1) / WG_THREADS);
  k_grad_p_update<<<nwg, WG_THREADS, 0, s>>>(g, p, dt, vel);
}

}  // namespace cup3d
