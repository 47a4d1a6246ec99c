// Learned warm start: a 3-layer 3-D CNN (2 -> 16 -> 16 -> 1 channels, 3^3
// kernels, GELU) applied per 8^3 block with a one-cell halo, trained and run
// by hand-written HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "poisson_internal.h"

namespace cup3d {

struct CnnState {
  void* d_w = nullptr;      // FP32 master weights
  void* d_w16 = nullptr;    // FP16 copy used by the forward pass
  void* d_grad = nullptr;
  void* d_adam = nullptr;   // m, v
  void* d_act = nullptr;    // activations for training
  void* d_tmp = nullptr;
  long long act_cap = 0;
  int nb = 0;
  const int* d_nbr = nullptr;
  int adam_t = 0;
  bool have_weights = false;
};

int cnn_num_weights();
int cnn_init(CnnState& C, std::string& err);
int cnn_set_topology(CnnState& C, int nb, const int* d_nbr, std::string& err);
int cnn_set_weights(CnnState& C, const float* w, int nw, std::string& err);
int cnn_get_weights(CnnState& C, float* w, int nw, std::string& err);
int cnn_predict(CnnState& C, hipStream_t s, const DevGrid<double>& g, const double* d_rhs,
                const double* d_prev, double* d_out, std::string& err);
int cnn_train_step(CnnState& C, hipStream_t s, const DevGrid<double>& g, const double* d_rhs,
                   const double* d_prev, const double* d_target, int nsamples, float lr,
                   double* loss, std::string& err);
// Replace the guess in d_x by the net's prediction when its residual is smaller.
int cnn_warm_start(CnnState& C, hipStream_t s, const DevGrid<double>& g,
                   const KrylovVectors<double>& v, SolverState* st, double* part,
                   const double* d_rhs, double* d_x, int bmean, int origin_block, int* used,
                   double* resid, std::string& err);
void cnn_free(CnnState& C);

}  // namespace cup3d
