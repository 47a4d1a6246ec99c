// Implicit diffusion kernels (diffusion.hip): the DST block preconditioner
// and the velocity gather/scatter of a diffusion step.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"
#include "poisson_internal.h"

namespace cup3d {

// Table of the DST block inverse (DST_TABLE doubles, device-resident for the
// kernels): S1[k][i] = sqrt(2/9) sin(pi (k+1)(i+1) / 9) row-major, then the
// eigenvalues l_k = 2 cos(pi (k+1) / 9) - 2 of the zero-padded 1-D second
// difference.
constexpr int DST_TABLE = BS * BS + BS;
void dst_tables(double* tab, float* tab32);  // FP64 table and its FP32 rounding

// mode PREC_P / PREC_S / PREC_PLAIN as launch_precond64; h per block,
// sigma = 1 / (nu dt).
void launch_precond_dst(hipStream_t s, int mode, int nb, const double* tab, const double* h,
                        double sigma, const KrylovVectors<double>& v, SolverState* st);
// The same in FP32 for the mixed forms' inner solve (sigma = 0: the pressure
// equation's block inverse h^-1 P_inv).
void launch_precond_dst32(hipStream_t s, int mode, int nb, const float* tab, const float* h,
                          const KrylovVectors<float>& v, SolverState* st);
// u = vel[:, comp], a = -h^3 sigma vel_old[:, comp] (vel, vel_old: [n][3]).
void launch_diffusion_gather(hipStream_t s, long long n, int comp, const double* h3, double sigma,
                             const double* vel, const double* vel_old, double* u, double* a);
// vel[:, comp] += x.
void launch_diffusion_scatter(hipStream_t s, long long n, int comp, const double* x,
                              double* vel);

}  // namespace cup3d
