// Learned warm start: a 3-layer 3-D CNN (2 -> 16 -> 16 -> 1 channels, 3^3
// kernels, GELU) applied per 8^3 block with a one-cell halo, trained and run
// by hand-written HIP kernels (cnn.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "poisson_internal.h"

namespace cup3d {

constexpr int CNN_C = 16;                                   // hidden channels
constexpr int CNN_NW1 = CNN_C * 2 * 27, CNN_NW2 = CNN_C * CNN_C * 27, CNN_NW3 = CNN_C * 27;
constexpr int CNN_OB1 = CNN_NW1, CNN_OW2 = CNN_OB1 + CNN_C, CNN_OB2 = CNN_OW2 + CNN_NW2;
constexpr int CNN_OW3 = CNN_OB2 + CNN_C, CNN_OB3 = CNN_OW3 + CNN_NW3;
constexpr int CNN_NW = CNN_OB3 + 1;                         // 8241
// FP16 fragment-packed weights: layer 1 7 k-steps, layers 2 and 3 14 each
constexpr int CNN_KT1 = 7, CNN_KT2 = 14, CNN_KT3 = 14;
constexpr int CNN_PK = (CNN_KT1 + CNN_KT2 + CNN_KT3) * 64 * 8;

struct CnnSample {
  float* rhs_h = nullptr;  // rhs / h_b        [n]
  float* prev = nullptr;   // p_prev           [n]
  float* target = nullptr; // (p - p_prev) / s [n]
  float s = 1.f;           // RMS(p_prev)
};

struct CnnState {
  float* d_w = nullptr;      // FP32 master weights [CNN_NW]
  float* d_adam = nullptr;   // m, v [2 * CNN_NW]
  void* d_wpk = nullptr;     // FP16 packed weights for inference
  double* d_scal = nullptr;  // device scalars: [0] = s of the last predict
  unsigned* d_counter = nullptr;
  double* d_part = nullptr;
  float* d_scratch = nullptr;  // training activations / gradients
  long long scratch_cap = 0;
  int* d_items = nullptr;
  int items_cap = 0;
  double* d_pred = nullptr;   // prediction buffer for the warm start
  int nb = 0;
  long long n = 0;
  const int* d_nbr = nullptr;
  const double* d_h = nullptr;
  int adam_t = 0;
  bool have_weights = false;
  std::vector<CnnSample> samples;
};

int cnn_num_weights();
int cnn_init(CnnState& C, std::string& err);
int cnn_set_topology(CnnState& C, int nb, const int* d_nbr, const double* d_h, std::string& err);
int cnn_set_weights(CnnState& C, const float* w, int nw, std::string& err);
int cnn_get_weights(CnnState& C, float* w, int nw, std::string& err);
int cnn_predict(CnnState& C, hipStream_t s, const double* d_rhs, const double* d_prev,
                double* d_out, std::string& err);
int cnn_add_sample(CnnState& C, hipStream_t s, const double* d_rhs, const double* d_prev,
                   const double* d_p, std::string& err);
void cnn_clear_samples(CnnState& C);
int cnn_train(CnnState& C, hipStream_t s, int steps, int batch, float lr, uint64_t seed,
              double* loss, std::string& err);
int cnn_eval(CnnState& C, hipStream_t s, int k, double* loss, std::string& err);
// Full-batch FP32 gradient of the loss over every block of sample k.
int cnn_gradient(CnnState& C, hipStream_t s, int k, float* grad, double* loss, std::string& err);
// Replace the guess in d_x by the net's prediction when its FP64 residual is
// smaller; reports the guess used and its residual norm.
int cnn_warm_start(CnnState& C, hipStream_t s, const DevGrid<double>& g,
                   const KrylovVectors<double>& v, SolverState* st, SolverState* h_st,
                   double* part, const double* d_rhs, double* d_x, int* used, double* resid,
                   std::string& err);
void cnn_free(CnnState& C);

}  // namespace cup3d
