// Mixed-precision solve: FP16/BF16 block-Jacobi GEMM with FP32 accumulation,
// FP32 Krylov vectors, FP64 iterative refinement.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cup3d_poisson.h"
#include "poisson_internal.h"

namespace cup3d {

struct MixedWorkspace {
  void* d_P16 = nullptr;   // FP16 P_inv, scaled, fragment-packed
  void* d_Pbf = nullptr;   // BF16 P_inv, fragment-packed
  bool split = true;       // hi/lo split input, two MFMA passes (false: one pass)
  double p16_scale = 1.0;  // P16 = P * p16_scale
  void* d_vec32 = nullptr; // FP32 Krylov vectors
  void* d_res64 = nullptr; // FP64 refinement residual
  long long n = 0;
  long long cap_n = 0;     // cells the vector buffers hold
  int nb = 0;
  int timing_precision = 1;
  KrylovVectors<float> v32{};
  // cfg.precond = DST: the FP32 inner solve applies the exact block inverse
  // by sine transforms (table dst32, DST_TABLE floats) instead of the
  // FP16/BF16 GEMM
  bool use_dst = false;
  const float* dst32 = nullptr;
  // in-solve kernel timing: HIP events around the main kernels of K1 .. K5
  // (the finalize kernels excluded) of the middle iteration of every polled
  // batch, summed while prof is set
  bool prof = false;
  hipEvent_t pev[10] = {};
  double prof_ms[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  long long prof_n = 0;
};

int mixed_init_pinv(MixedWorkspace& M, const std::vector<double>& P, std::string& err);
int mixed_set_topology(MixedWorkspace& M, int nb, long long n, std::string& err);
int mixed_solve(MixedWorkspace& M, hipStream_t s, const cup3d_poisson_config& cfg,
                const DevGrid<double>& g64, const DevGrid<float>& g32,
                const KrylovVectors<double>& v64, SolverState* st, SolverState* h_st,
                double* part, int origin_block, const double* d_rhs, double* d_x,
                const double* d_x0, cup3d_solve_report* rep, std::string& err);
// y = A x with the FP32 operator of the inner solve (x, y FP64 device vectors).
int mixed_apply_operator32(MixedWorkspace& M, hipStream_t s, const DevGrid<float>& g32,
                           SolverState* st, double* part, const double* d_x, double* d_y,
                           std::string& err);
int mixed_apply_precond(MixedWorkspace& M, hipStream_t s, int precision, const DevGrid<float>& g32,
                        const double* d_p, double* d_z, std::string& err);
int mixed_time_kernel(MixedWorkspace& M, hipStream_t s, int which, int reps,
                      const DevGrid<float>& g32, SolverState* st, double* part, hipEvent_t e0,
                      hipEvent_t e1, double* avg_ms, double* bytes, double* flops,
                      std::string& err);
void mixed_free(MixedWorkspace& M);

}  // namespace cup3d
