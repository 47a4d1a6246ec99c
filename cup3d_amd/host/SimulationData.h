// The part of CubismUP_3D's SimulationData and Cubism's block grid that the
// pressure solver reads (source/SimulationData.h:97-110, Definitions.h:38-50,
// 524-531; BlockInfo as used at PoissonSolverExp.cpp:375-472). The Cubism
// library itself is not part of this package; a caller fills these from its
// own grid.
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace cubismup3d {

enum BCflag { freespace, periodic, wall };  // Definitions.h:38

constexpr int kBlockSize = 8;

// One 8x8x8 block of doubles, [iz][iy][ix] (cubism::GridBlock<8,3,ScalarElement>).
struct ScalarBlock {
  double data[kBlockSize][kBlockSize][kBlockSize];
  double& operator()(int ix, int iy, int iz) { return data[iz][iy][ix]; }
  const double& operator()(int ix, int iy, int iz) const { return data[iz][iy][ix]; }
};

// The BlockInfo fields the solver uses.
struct BlockInfo {
  long long blockID = 0;
  int level = 0;
  int index[3] = {0, 0, 0};
  double h = 0;
  void* ptrBlock = nullptr;  // ScalarBlock*
};

struct ScalarGrid {
  std::vector<BlockInfo> infos;
  bool UpdateFluxCorrection = true;  // set when the mesh changed (Simulation.cpp:146)
};

struct SimulationData {
  std::string poissonSolver = "hip_iterative";  // -poissonSolver
  double PoissonErrorTol = 1e-6;                // -poissonTol (SimulationData.cpp:85)
  double PoissonErrorTolRel = 1e-4;             // -poissonTolRel (:86)
  int bMeanConstraint = 1;                      // -bMeanConstraint (:87)
  BCflag BCx_flag = freespace, BCy_flag = freespace, BCz_flag = freespace;  // -BC_x/y/z
  int bpdx = 1, bpdy = 1, bpdz = 1;             // blocks per direction at level 0
  int levelMax = 1;
  double h0 = 1.0;                              // level-0 cell size
  int device = 0;
  ScalarGrid* lhs = nullptr;   // rhs of the Poisson equation
  ScalarGrid* pres = nullptr;  // guess on input, pressure on output
  std::vector<BlockInfo>& lhsInfo() { return lhs->infos; }
  std::vector<BlockInfo>& presInfo() { return pres->infos; }
};

}  // namespace cubismup3d
