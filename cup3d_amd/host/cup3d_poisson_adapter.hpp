// PoissonSolverBase adapter over libcup3d_poisson's C ABI, header-only.
//
// Replaces PoissonSolverExp (source/poisson/PoissonSolverExp.{h,cpp}) behind
// the reference's own interface. The adapter is a template over the caller's
// types, so it compiles inside CubismUP_3D against the reference's
// declarations and defines none of its names:
//   Base  the caller's solver interface with `virtual void solve() = 0`
//         (cubismup3d::PoissonSolverBase, source/poisson/PoissonSolverBase.h:10-17)
//   Sim   the caller's simulation data (cubismup3d::SimulationData,
//         source/SimulationData.h:37-110); the members read are
//           poissonSolver, PoissonErrorTol, PoissonErrorTolRel, bMeanConstraint,
//           BCx_flag / BCy_flag / BCz_flag, bpdx / bpdy / bpdz, levelMax,
//           hmax (the level-0 spacing, SimulationData.cpp:162-163),
//           lhsInfo(), presInfo() and pres->UpdateFluxCorrection,
//         and of every BlockInfo: level, index[3], ptrBlock (an 8x8x8 block of
//         doubles, [iz][iy][ix]).
// Everything here lives in namespace cup3d_hip. A standalone stand-in for the
// two types (for callers without Cubism, and for this package's own example)
// is in cup3d_poisson_standalone.hpp.
//
// Registration in the reference factory (source/poisson/PoissonSolverBase.cpp:10-32):
//   if (auto p = cup3d_hip::makePoissonSolver<PoissonSolverBase>(s)) return p;
// answers the names of this library and returns nullptr for every other name.
#pragma once

#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cup3d_poisson.h"

namespace cup3d_hip {

// Solver names of this library and the arithmetic each selects; -1 for any
// other name. "cuda_iterative" (PoissonSolverBase.cpp:16) is answered by the
// FP64 drop-in so launch scripts of the reference keep working.
inline int precisionFor(const std::string& name) {
  if (name == "hip_iterative" || name == "cuda_iterative") return CUP3D_PREC_FP64;
  if (name == "hip_iterative_mixed") return CUP3D_PREC_MIXED_FP16;
  if (name == "hip_iterative_mixed_bf16") return CUP3D_PREC_MIXED_BF16;
  return -1;
}

template <class Base, class Sim>
class PoissonSolverHIP final : public Base {
 public:
  // PoissonSolverExp ctor (PoissonSolverExp.cpp:37-111): configuration from
  // the simulation data; the operator is built on the first solve.
  PoissonSolverHIP(Sim& s, int precision, int device = 0) : sim(s) {
    cup3d_poisson_config c{};
    c.bpd[0] = s.bpdx;
    c.bpd[1] = s.bpdy;
    c.bpd[2] = s.bpdz;
    c.level_max = s.levelMax;
    c.bc[0] = static_cast<int>(s.BCx_flag);  // BCflag order: freespace, periodic, wall
    c.bc[1] = static_cast<int>(s.BCy_flag);
    c.bc[2] = static_cast<int>(s.BCz_flag);
    c.h0 = static_cast<double>(s.hmax);
    c.tol_abs = static_cast<double>(s.PoissonErrorTol);     // PoissonSolverExp.cpp:433
    c.tol_rel = static_cast<double>(s.PoissonErrorTolRel);  // :434
    c.tol_rhs = 0.0;
    c.bmean = s.bMeanConstraint;
    c.max_iter = 1000;     // BiCGSTAB.cu:474
    c.max_restarts = 100;  // PoissonSolverExp.cpp:435
    c.precision = precision;
    c.guess = CUP3D_GUESS_GIVEN;
    c.device = device;
    if (cup3d_poisson_create(&c, &handle_) != 0) raise(nullptr, "create");
  }
  ~PoissonSolverHIP() override { cup3d_poisson_destroy(handle_); }
  PoissonSolverHIP(const PoissonSolverHIP&) = delete;
  PoissonSolverHIP& operator=(const PoissonSolverHIP&) = delete;

  // PoissonSolverExp::solve (PoissonSolverExp.cpp:420-473): rebuild the
  // operator when the mesh changed, gather rhs and guess from the blocks,
  // solve, scatter the pressure into sim.pres.
  void solve() override {
    auto& rhsInfo = sim.lhsInfo();
    auto& xInfo = sim.presInfo();
    const size_t nb = rhsInfo.size();
    if (xInfo.size() != nb) throw std::runtime_error("Poisson solver: lhs/pres block counts differ");
    if (sim.pres->UpdateFluxCorrection) {  // mesh changed (PoissonSolverExp.cpp:437-442)
      std::vector<cup3d_block_desc> d(nb);
      for (size_t i = 0; i < nb; ++i) {
        d[i].level = rhsInfo[i].level;
        for (int a = 0; a < 3; ++a) d[i].index[a] = rhsInfo[i].index[a];
      }
      if (cup3d_poisson_set_topology(handle_, (int64_t)nb, d.data()) != 0) raise(handle_, "topology");
      sim.pres->UpdateFluxCorrection = false;
    }
    rhs_ptrs_.resize(nb);
    x_ptrs_.resize(nb);
    for (size_t i = 0; i < nb; ++i) {
      rhs_ptrs_[i] = static_cast<const double*>(rhsInfo[i].ptrBlock);
      x_ptrs_[i] = static_cast<double*>(xInfo[i].ptrBlock);
    }
    const int st = cup3d_poisson_solve(handle_, rhs_ptrs_.data(), x_ptrs_.data(), nullptr,
                                       nullptr, &report_);
    if (st < 0) raise(handle_, "solve");
    // like BiCGSTAB.cu:612-621: non-convergence is reported, not raised
    if (st != CUP3D_CONVERGED)
      std::cout << "  [BiCGSTAB]: Iteration " << report_.iterations
                << ". Error norm = " << report_.err_final << " (status " << st << ")\n";
  }

  const cup3d_solve_report& lastReport() const { return report_; }
  cup3d_poisson_t handle() const { return handle_; }

 private:
  [[noreturn]] static void raise(cup3d_poisson_t h, const std::string& what) {
    throw std::runtime_error("Poisson solver: " + what + ": " + cup3d_poisson_last_error(h));
  }

  Sim& sim;
  cup3d_poisson_t handle_ = nullptr;
  cup3d_solve_report report_{};
  std::vector<const double*> rhs_ptrs_;
  std::vector<double*> x_ptrs_;
};

// The solver for s.poissonSolver if it names one of this library's solvers,
// else nullptr (the caller's factory goes on with its own names).
template <class Base, class Sim>
std::shared_ptr<Base> makePoissonSolver(Sim& s, int device = 0) {
  const int p = precisionFor(s.poissonSolver);
  if (p < 0) return nullptr;
  return std::make_shared<PoissonSolverHIP<Base, Sim>>(s, p, device);
}

}  // namespace cup3d_hip
