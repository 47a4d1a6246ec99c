// Drives the solver the way PressureProjection does (PressureProjection.cpp:
// 325-402): construct through makePoissonSolver, put the rhs in sim.lhs and
// the guess in sim.pres, call solve(), read the pressure from sim.pres.
// The rhs is one periodic Fourier mode on a uniform grid, whose discrete
// solution is known exactly: x = h^2 f / sum_d (2 cos(k_d h) - 2).
// Prints one JSON line per solve.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "cup3d_poisson_standalone.hpp"

using namespace cup3d_hip::standalone;

int main(int argc, char** argv) {
  const std::string solver = argc > 1 ? argv[1] : "hip_iterative";
  const int bp = argc > 2 ? std::atoi(argv[2]) : 4;
  SimulationData sim;
  sim.poissonSolver = solver;
  sim.PoissonErrorTol = 0.0;
  sim.PoissonErrorTolRel = 1e-7;
  sim.bMeanConstraint = 1;
  sim.BCx_flag = sim.BCy_flag = sim.BCz_flag = periodic;
  sim.bpdx = sim.bpdy = sim.bpdz = bp;
  sim.hmax = 1.0 / (8 * bp);
  ScalarGrid lhs, pres;
  sim.lhs = &lhs;
  sim.pres = &pres;
  const int nb = bp * bp * bp;
  std::vector<ScalarBlock> rb(nb), pb(nb);
  const double h = sim.hmax, k[3] = {2 * M_PI * 1, 2 * M_PI * 2, 2 * M_PI * 3};
  double lam = 0;
  for (int d = 0; d < 3; ++d) lam += 2 * std::cos(k[d] * h) - 2;
  int id = 0;
  for (int bz = 0; bz < bp; ++bz)
    for (int by = 0; by < bp; ++by)
      for (int bx = 0; bx < bp; ++bx, ++id) {
        BlockInfo info;
        info.blockID = id;
        info.level = 0;
        info.index[0] = bx;
        info.index[1] = by;
        info.index[2] = bz;
        info.h = h;
        info.ptrBlock = &rb[id];
        lhs.infos.push_back(info);
        info.ptrBlock = &pb[id];
        pres.infos.push_back(info);
        for (int iz = 0; iz < 8; ++iz)
          for (int iy = 0; iy < 8; ++iy)
            for (int ix = 0; ix < 8; ++ix) {
              const double x = (bx * 8 + ix + 0.5) * h, y = (by * 8 + iy + 0.5) * h,
                           z = (bz * 8 + iz + 0.5) * h;
              rb[id](ix, iy, iz) = h * h * h * std::cos(k[0] * x + k[1] * y + k[2] * z);
              pb[id](ix, iy, iz) = 0.0;
            }
      }
  auto S = makePoissonSolver(sim);
  for (int rep = 0; rep < 2; ++rep) {
    if (rep == 1)  // second step: warm start from the first solution
      for (auto& b : pb) b(1, 1, 1) += 1e-3;
    S->solve();
    double err = 0, ref = 0;
    id = 0;
    for (int bz = 0; bz < bp; ++bz)
      for (int by = 0; by < bp; ++by)
        for (int bx = 0; bx < bp; ++bx, ++id)
          for (int iz = 0; iz < 8; ++iz)
            for (int iy = 0; iy < 8; ++iy)
              for (int ix = 0; ix < 8; ++ix) {
                const double x = (bx * 8 + ix + 0.5) * h, y = (by * 8 + iy + 0.5) * h,
                             z = (bz * 8 + iz + 0.5) * h;
                const double ex = h * h * std::cos(k[0] * x + k[1] * y + k[2] * z) / lam;
                err = std::fmax(err, std::fabs(pb[id](ix, iy, iz) - ex));
                ref = std::fmax(ref, std::fabs(ex));
              }
    const auto& r = static_cast<PoissonSolverHIP&>(*S).lastReport();
    std::printf("{\"solver\": \"%s\", \"solve\": %d, \"status\": %d, \"iterations\": %d, "
                "\"max_rel_error\": %.3e, \"true_rel_resid\": %.3e}\n",
                solver.c_str(), rep, r.status, r.iterations, err / ref, r.true_rel_resid);
  }
  try {
    sim.poissonSolver = "no_such_solver";
    makePoissonSolver(sim);
    std::printf("{\"factory_error\": false}\n");
  } catch (const std::invalid_argument&) {
    std::printf("{\"factory_error\": true}\n");
  }
  return 0;
}
