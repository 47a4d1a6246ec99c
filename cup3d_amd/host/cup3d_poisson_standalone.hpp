// Standalone stand-ins for the caller types the adapter reads, for programs
// that use the solver without CubismUP_3D (this package's example caller and
// tests). They carry only what the adapter reads, with the reference's member
// names (source/SimulationData.h:37-110, source/Definitions.h:38-50, 524-531;
// BlockInfo as used at PoissonSolverExp.cpp:375-472), in their own namespace
// so they never meet the reference's types. Inside CubismUP_3D the adapter is
// instantiated on the reference's own types instead.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cup3d_poisson_adapter.hpp"

namespace cup3d_hip {
namespace standalone {

enum BCflag { freespace, periodic, wall };

constexpr int kBlockSize = 8;

// One 8x8x8 block of doubles, [iz][iy][ix].
struct ScalarBlock {
  double data[kBlockSize][kBlockSize][kBlockSize];
  double& operator()(int ix, int iy, int iz) { return data[iz][iy][ix]; }
  const double& operator()(int ix, int iy, int iz) const { return data[iz][iy][ix]; }
};

struct BlockInfo {
  long long blockID = 0;
  int level = 0;
  int index[3] = {0, 0, 0};
  double h = 0;
  void* ptrBlock = nullptr;  // ScalarBlock*
};

struct ScalarGrid {
  std::vector<BlockInfo> infos;
  bool UpdateFluxCorrection = true;  // set when the mesh changed
};

struct SimulationData {
  std::string poissonSolver = "hip_iterative";
  double PoissonErrorTol = 1e-6;
  double PoissonErrorTolRel = 1e-4;
  int bMeanConstraint = 1;
  BCflag BCx_flag = freespace, BCy_flag = freespace, BCz_flag = freespace;
  int bpdx = 1, bpdy = 1, bpdz = 1;
  int levelMax = 1;
  double hmax = 1.0;  // level-0 cell size
  ScalarGrid* lhs = nullptr;   // rhs of the Poisson equation
  ScalarGrid* pres = nullptr;  // guess on input, pressure on output
  std::vector<BlockInfo>& lhsInfo() const { return lhs->infos; }
  std::vector<BlockInfo>& presInfo() const { return pres->infos; }
};

class PoissonSolverBase {
 public:
  virtual ~PoissonSolverBase() = default;
  virtual void solve() = 0;
};

using PoissonSolverHIP = cup3d_hip::PoissonSolverHIP<PoissonSolverBase, SimulationData>;

// Factory with the reference's error behaviour for names it does not know
// (PoissonSolverBase.cpp:27-31: std::invalid_argument).
inline std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData& s) {
  if (auto p = cup3d_hip::makePoissonSolver<PoissonSolverBase>(s)) return p;
  throw std::invalid_argument("Poisson solver: \"" + s.poissonSolver + "\" unrecognized!");
}

}  // namespace standalone
}  // namespace cup3d_hip
