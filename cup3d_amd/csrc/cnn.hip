// Learned warm start (placeholder until the CNN kernels land).
#include "cnn.h"

namespace cup3d {
int cnn_num_weights() { return 2 * 16 * 27 + 16 + 16 * 16 * 27 + 16 + 16 * 27 + 1; }
int cnn_init(CnnState&, std::string&) { return 0; }
int cnn_set_topology(CnnState& C, int nb, const int* d_nbr, std::string&) {
  C.nb = nb;
  C.d_nbr = d_nbr;
  return 0;
}
int cnn_set_weights(CnnState&, const float*, int, std::string& err) { err = "cnn not built yet"; return -1; }
int cnn_get_weights(CnnState&, float*, int, std::string& err) { err = "cnn not built yet"; return -1; }
int cnn_predict(CnnState&, hipStream_t, const DevGrid<double>&, const double*, const double*,
                double*, std::string& err) { err = "cnn not built yet"; return -1; }
int cnn_train_step(CnnState&, hipStream_t, const DevGrid<double>&, const double*, const double*,
                   const double*, int, float, double*, std::string& err) { err = "cnn not built yet"; return -1; }
int cnn_warm_start(CnnState&, hipStream_t, const DevGrid<double>&, const KrylovVectors<double>&,
                   SolverState*, double*, const double*, double*, int, int, int*, double*,
                   std::string& err) { err = "cnn not built yet"; return -1; }
void cnn_free(CnnState&) {}
}  // namespace cup3d
