// PoissonSolverBase implementation on top of libcup3d_poisson's C ABI.
// Replaces PoissonSolverExp (source/poisson/PoissonSolverExp.{h,cpp}).
#pragma once

#include <string>
#include <vector>

#include "PoissonSolverBase.h"
#include "SimulationData.h"
#include "cup3d_poisson.h"

namespace cubismup3d {

class PoissonSolverHIP : public PoissonSolverBase {
 public:
  PoissonSolverHIP(SimulationData& s, int precision);
  ~PoissonSolverHIP() override;
  PoissonSolverHIP(const PoissonSolverHIP&) = delete;
  PoissonSolverHIP& operator=(const PoissonSolverHIP&) = delete;

  // PoissonSolverExp::solve (PoissonSolverExp.cpp:420-473): rebuild the
  // operator when the mesh changed, gather rhs and guess from the blocks,
  // solve, scatter the pressure into sim.pres.
  void solve() override;

  const cup3d_solve_report& lastReport() const { return report_; }

 private:
  SimulationData& sim;
  cup3d_poisson_t handle_ = nullptr;
  cup3d_solve_report report_{};
  std::vector<const double*> rhs_ptrs_;
  std::vector<double*> x_ptrs_;
};

}  // namespace cubismup3d
