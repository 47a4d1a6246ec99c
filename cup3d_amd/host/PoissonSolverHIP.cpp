#include "PoissonSolverHIP.h"

#include <iostream>
#include <stdexcept>

namespace cubismup3d {

namespace {
[[noreturn]] void raise(cup3d_poisson_t h, const std::string& what) {
  throw std::runtime_error("Poisson solver: " + what + ": " + cup3d_poisson_last_error(h));
}
}  // namespace

PoissonSolverHIP::PoissonSolverHIP(SimulationData& s, int precision) : sim(s) {
  cup3d_poisson_config c{};
  c.bpd[0] = s.bpdx;
  c.bpd[1] = s.bpdy;
  c.bpd[2] = s.bpdz;
  c.level_max = s.levelMax;
  c.bc[0] = s.BCx_flag;
  c.bc[1] = s.BCy_flag;
  c.bc[2] = s.BCz_flag;
  c.h0 = s.h0;
  c.tol_abs = s.PoissonErrorTol;      // PoissonSolverExp.cpp:433
  c.tol_rel = s.PoissonErrorTolRel;   // :434
  c.tol_rhs = 0.0;
  c.bmean = s.bMeanConstraint;
  c.max_iter = 1000;                  // BiCGSTAB.cu:474
  c.max_restarts = 100;               // PoissonSolverExp.cpp:435
  c.precision = precision;
  c.guess = CUP3D_GUESS_GIVEN;
  c.device = s.device;
  if (cup3d_poisson_create(&c, &handle_) != 0) raise(nullptr, "create");
}

PoissonSolverHIP::~PoissonSolverHIP() { cup3d_poisson_destroy(handle_); }

void PoissonSolverHIP::solve() {
  std::vector<BlockInfo>& rhsInfo = sim.lhsInfo();
  std::vector<BlockInfo>& xInfo = sim.presInfo();
  const size_t nb = rhsInfo.size();
  if (xInfo.size() != nb) throw std::runtime_error("Poisson solver: lhs/pres block counts differ");
  if (sim.pres->UpdateFluxCorrection) {  // mesh changed (PoissonSolverExp.cpp:437-442)
    sim.pres->UpdateFluxCorrection = false;
    std::vector<cup3d_block_desc> d(nb);
    for (size_t i = 0; i < nb; ++i) {
      d[i].level = rhsInfo[i].level;
      for (int a = 0; a < 3; ++a) d[i].index[a] = rhsInfo[i].index[a];
    }
    if (cup3d_poisson_set_topology(handle_, (int64_t)nb, d.data()) != 0) raise(handle_, "topology");
  }
  rhs_ptrs_.resize(nb);
  x_ptrs_.resize(nb);
  for (size_t i = 0; i < nb; ++i) {
    rhs_ptrs_[i] = static_cast<const double*>(rhsInfo[i].ptrBlock);
    x_ptrs_[i] = static_cast<double*>(xInfo[i].ptrBlock);
  }
  const int st = cup3d_poisson_solve(handle_, rhs_ptrs_.data(), x_ptrs_.data(), nullptr, nullptr,
                                     &report_);
  if (st < 0) raise(handle_, "solve");
  // like BiCGSTAB.cu:612-621: non-convergence is reported, not raised
  if (st != CUP3D_CONVERGED)
    std::cout << "  [BiCGSTAB]: Iteration " << report_.iterations << ". Error norm = "
              << report_.err_final << " (status " << st << ")\n";
}

std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData& s) {
  if (s.poissonSolver == "hip_iterative" || s.poissonSolver == "cuda_iterative")
    return std::make_shared<PoissonSolverHIP>(s, CUP3D_PREC_FP64);
  if (s.poissonSolver == "hip_iterative_mixed")
    return std::make_shared<PoissonSolverHIP>(s, CUP3D_PREC_MIXED_FP16);
  if (s.poissonSolver == "hip_iterative_mixed_bf16")
    return std::make_shared<PoissonSolverHIP>(s, CUP3D_PREC_MIXED_BF16);
  if (s.poissonSolver == "iterative")
    throw std::runtime_error("Poisson solver: \"iterative\" is the reference CPU solver; "
                             "this library provides \"hip_iterative\"");
  throw std::invalid_argument("Poisson solver: \"" + s.poissonSolver + "\" unrecognized!");
}

}  // namespace cubismup3d
