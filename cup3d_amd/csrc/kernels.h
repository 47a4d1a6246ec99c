// Kernel launch interface of libcup3d_poisson (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "poisson_internal.h"

namespace cup3d {

enum SpmvMode { SPMV_PLAIN = 0, SPMV_INIT = 1, SPMV_DOT1 = 2, SPMV_DOT2 = 3, SPMV_RESID = 4 };
enum PrecMode { PREC_PLAIN = 0, PREC_P = 1, PREC_S = 2 };

void launch_getvec64(hipStream_t s, long long n, long long zero_row, const double* rhs,
                     const double* x0, double* b, double* x, SolverState* st, double* part,
                     double tol_b);
void launch_spmv64(hipStream_t s, int mode, const DevGrid<double>& g, const double* x, double* y,
                   const double* a, double* c, double* c2, SolverState* st, double* part);
void launch_precond64(hipStream_t s, int mode, int nb, const double* Ppk, const double* invh,
                      const KrylovVectors<double>& v, SolverState* st);
void launch_update64(hipStream_t s, long long n, const KrylovVectors<double>& v, SolverState* st,
                     double* part);
void launch_copy_opt64(hipStream_t s, long long n, const KrylovVectors<double>& v,
                       const SolverState* st, double* out);

// The caller around the solve (projection.hip): the rhs of the pressure
// equation from the velocity (and an obstacle's udef / chi, both or neither
// null), and the pressure-gradient update of the velocity ([n][3] doubles).
void launch_pressure_rhs(hipStream_t s, const DevGrid<double>& g, const double* vel,
                         const double* udef, const double* chi, double dt, double* out);
void launch_grad_p_update(hipStream_t s, const DevGrid<double>& g, const double* p, double dt,
                          double* vel);

// x -= the volume-weighted mean of x (h3 per block, inv_vol = 1 / sum 512 h^3);
// the weighted sum is left in wsum[0].
void launch_remove_mean(hipStream_t s, long long n, double* x, const double* h3, double inv_vol,
                        double* part, double* wsum);
// b = a over n4 float4 (the copy-rate probe of the diagnostics).
void launch_stream_copy(hipStream_t s, long long n4, const void* a, void* b);

// Warm start from several guesses (device vectors of n doubles).
constexpr int HIST_MAX = 8;                // differences of previous pressures
constexpr int GUESS_MAXV = HIST_MAX + 2;   // r0 + the history directions + the net's
struct GuessVecs {
  const double* v[GUESS_MAXV];
};
// d = a - b, and d[row] = 0 when row >= 0 (the bMean = 1 / 3 cell that getVec
// zeroes): one direction of the warm start;
void launch_guess_direction(hipStream_t s, long long n, const double* a, const double* b,
                            long long row, double* d);
// x[row] = 0 (nothing for row < 0): the bMean = 1 / 3 cell that getVec zeroes;
void launch_zero_cell(hipStream_t s, double* x, long long row);
// dots = the upper triangle of the Gram matrix of r.v[0 .. nv), row by row,
// GUESS_MAXV (GUESS_MAXV + 1) / 2 values (those of absent vectors 0), in a
// fixed order; part holds that many partials per 2048 cells;
void launch_guess_gram(hipStream_t s, long long n, const GuessVecs& r, int nv, double* part,
                       double* dots);
// x = x0 + sum_{k < m} gt[k] (h[k-1] - h[k]) + gnet (pred - x0), h[-1] = x0,
// and x[row] = x0[row] (row >= 0); pred is not read when gnet = 0.
void launch_guess_blend(hipStream_t s, long long n, const double* x0, const GuessVecs& h,
                        const double* gt, int m, const double* pred, double gnet, long long row,
                        double* x);

}  // namespace cup3d
