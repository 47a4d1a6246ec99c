// Solver interface of CubismUP_3D's pressure projection, as the reference
// declares it (source/poisson/PoissonSolverBase.h:10-19): an argument-less
// solve() that reads sim.lhs (rhs) and sim.pres (guess) and writes sim.pres.
#pragma once

#include <memory>

namespace cubismup3d {

struct SimulationData;

class PoissonSolverBase {
 public:
  virtual ~PoissonSolverBase() = default;
  virtual void solve() = 0;
};

// String factory (PoissonSolverBase.cpp:30-52). Recognised names:
//   "hip_iterative" (alias "cuda_iterative")  FP64 drop-in on the GPU
//   "hip_iterative_mixed"                     FP16 GEMM / FP32 Krylov / FP64 refinement
//   "hip_iterative_mixed_bf16"                same with BF16
// Unknown names throw std::invalid_argument, like the reference.
std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData& s);

}  // namespace cubismup3d
