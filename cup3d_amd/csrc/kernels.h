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

}  // namespace cup3d
