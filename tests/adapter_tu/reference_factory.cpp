// A translation unit shaped like the reference's factory file
// (source/poisson/PoissonSolverBase.cpp:1-32): the reference's declarations
// come first (restated here with the members the adapter reads:
// source/poisson/PoissonSolverBase.h:10-19, source/SimulationData.h:37-110,
// source/Definitions.h:38, Cubism's BlockInfo and grid), then the adapter is
// included, then makePoissonSolver gains the branch INTEGRATION.md shows.
// It must compile without a redefinition and link against
// libcup3d_poisson.so with no undefined symbol.
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace cubism {
struct BlockInfo {
  long long blockID;
  int level;
  int index[3];
  double h;
  double origin[3];
  void* ptrBlock;
};
}  // namespace cubism

namespace cubismup3d {
using Real = double;
enum BCflag { freespace, periodic, wall, BCflagNone };
struct ScalarBlock;
struct ScalarGrid {
  typedef ScalarBlock BlockType;
  bool UpdateFluxCorrection = true;
  std::vector<cubism::BlockInfo> infos;
  std::vector<cubism::BlockInfo>& getBlocksInfo() { return infos; }
};

struct SimulationData;

class PoissonSolverBase {
 public:
  virtual ~PoissonSolverBase() = default;
  virtual void solve() = 0;

 protected:
  typedef typename ScalarGrid::BlockType BlockType;
};

std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData& s);

struct SimulationData {
  ScalarGrid* pres = nullptr;
  ScalarGrid* lhs = nullptr;
  inline std::vector<cubism::BlockInfo>& presInfo() const { return pres->getBlocksInfo(); }
  inline std::vector<cubism::BlockInfo>& lhsInfo() const { return lhs->getBlocksInfo(); }
  int bpdx, bpdy, bpdz;
  int levelMax;
  Real hmin, hmax;
  Real PoissonErrorTol;
  Real PoissonErrorTolRel;
  std::string poissonSolver;
  BCflag BCx_flag = freespace;
  BCflag BCy_flag = freespace;
  BCflag BCz_flag = freespace;
  int bMeanConstraint = 1;
};

class PoissonSolverAMR : public PoissonSolverBase {
 public:
  explicit PoissonSolverAMR(SimulationData&) {}
  void solve() override {}
};
}  // namespace cubismup3d

// ---- the two lines a maintainer adds to PoissonSolverBase.cpp ----
#include "cup3d_poisson_adapter.hpp"

namespace cubismup3d {
std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData& s) {
  if (auto p = cup3d_hip::makePoissonSolver<PoissonSolverBase>(s)) return p;
  // ---- the reference's own branches follow unchanged ----
  if (s.poissonSolver == "iterative") return std::make_shared<PoissonSolverAMR>(s);
  throw std::invalid_argument("Poisson solver: \"" + s.poissonSolver + "\" unrecognized!");
}
}  // namespace cubismup3d

// The caller side: PressureProjection's constructor and operator()
// (source/operators/PressureProjection.cpp:325-329, 402).
extern "C" int reference_style_projection(int bpd) {
  using namespace cubismup3d;
  SimulationData sim;
  ScalarGrid lhs, pres;
  sim.lhs = &lhs;
  sim.pres = &pres;
  sim.bpdx = sim.bpdy = sim.bpdz = bpd;
  sim.levelMax = 1;
  sim.hmin = sim.hmax = 1.0 / (8 * bpd);
  sim.PoissonErrorTol = 0.0;
  sim.PoissonErrorTolRel = 1e-7;
  sim.poissonSolver = "hip_iterative";
  sim.BCx_flag = sim.BCy_flag = sim.BCz_flag = periodic;
  std::shared_ptr<PoissonSolverBase> solver = makePoissonSolver(sim);
  solver->solve();
  return 0;
}
