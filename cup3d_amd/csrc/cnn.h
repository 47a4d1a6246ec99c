// Learned warm start: a 3-layer 3-D CNN (2 -> 16 -> 16 -> 1 channels, 3^3
// kernels, GELU) applied per 8^3 block with a one-cell halo, trained and run
// by hand-written HIP kernels (cnn.hip). It predicts the pressure increment
// of a time step from the residual of the previous pressure.
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
  float* res_h = nullptr;   // (b - A p_prev) / h_b  [n]
  float* prev = nullptr;    // p_prev               [n]
  float* target = nullptr;  // (p - p_prev) / s_r   [n]
  float s_r = 1.f;          // RMS((b - A p_prev) / h_b)
  float s_p = 1.f;          // RMS(p_prev)
};

struct CnnState {
  float* d_w = nullptr;      // FP32 master weights [CNN_NW]
  float* d_adam = nullptr;   // m, v [2 * CNN_NW]
  void* d_wpk = nullptr;     // FP16 packed weights for inference
  double* d_scal = nullptr;  // device scalars: [0] = s_r, [1] = s_p of the last predict
  double* d_part = nullptr;
  float* d_scratch = nullptr;  // training activations / gradients
  long long scratch_cap = 0;
  int* d_items = nullptr;
  int items_cap = 0;
  double* d_pred = nullptr;   // prediction buffer for the warm start
  int cap_nb = 0;             // blocks d_part and d_pred hold
  int nb = 0;
  long long n = 0;
  const int* d_nbr = nullptr;
  const double* d_h = nullptr;
  int adam_t = 0;
  float res_weight = 0.f;  // weight of the local-residual loss term of the last training
  bool have_weights = false;
  std::vector<CnnSample> samples;
};

int cnn_num_weights();
int cnn_init(CnnState& C, std::string& err);
int cnn_set_topology(CnnState& C, int nb, const int* d_nbr, const double* d_h, std::string& err);
int cnn_set_weights(CnnState& C, const float* w, int nw, std::string& err);
int cnn_get_weights(CnnState& C, float* w, int nw, std::string& err);
// d_out = p_prev + s_r * net(res / (h s_r), p_prev / s_p), res = b - A p_prev.
int cnn_predict(CnnState& C, hipStream_t s, const double* d_res, const double* d_prev,
                double* d_out, std::string& err);
// Training triple: the residual of p_prev, p_prev, and the solution p.
int cnn_add_sample(CnnState& C, hipStream_t s, const double* d_res, const double* d_prev,
                   const double* d_p, std::string& err);
void cnn_clear_samples(CnnState& C);
// Loss per item: 0.5 mean((out - target)^2) + res_weight 0.5 mean(rho^2),
// rho = lap7(out) - res / (h s_r) with the target closing the stencil
// outside the block (the discrete increment equation A delta = res).
int cnn_train(CnnState& C, hipStream_t s, int steps, int batch, float lr, float res_weight,
              uint64_t seed, double* loss, std::string& err);
int cnn_eval(CnnState& C, hipStream_t s, int k, double* loss, std::string& err);
// Full-batch FP32 gradient of the loss over every block of sample k.
int cnn_gradient(CnnState& C, hipStream_t s, int k, float res_weight, float* grad, double* loss,
                 std::string& err);
void cnn_free(CnnState& C);

}  // namespace cup3d
