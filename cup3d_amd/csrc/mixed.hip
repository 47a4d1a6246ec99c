// Mixed-precision solve (placeholder until the FP32/FP16 kernels land).
#include "mixed.h"

namespace cup3d {
int mixed_init_pinv(MixedWorkspace&, const std::vector<double>&, std::string&) { return 0; }
int mixed_set_topology(MixedWorkspace& M, int nb, long long n, std::string&) {
  M.nb = nb;
  M.n = n;
  return 0;
}
int mixed_solve(MixedWorkspace&, hipStream_t, const cup3d_poisson_config&, const DevGrid<double>&,
                const DevGrid<float>&, const KrylovVectors<double>&, SolverState*, SolverState*,
                double*, int, const double*, double*, const double*, cup3d_solve_report*,
                std::string& err) {
  err = "mixed precision not built yet";
  return -1;
}
int mixed_apply_precond(MixedWorkspace&, hipStream_t, int, const DevGrid<float>&, const double*,
                        double*, std::string& err) {
  err = "mixed precision not built yet";
  return -1;
}
int mixed_time_kernel(MixedWorkspace&, hipStream_t, int, int, const DevGrid<float>&, SolverState*,
                      double*, hipEvent_t, hipEvent_t, double*, double*, double*,
                      std::string& err) {
  err = "mixed precision not built yet";
  return -1;
}
void mixed_free(MixedWorkspace&) {}
}  // namespace cup3d
