// CPU ORACLE — TEST INFRASTRUCTURE ONLY.
//
// This file is a plain FP64 CPU restatement of the reference pressure-Poisson
// path of CubismUP_3D. It is the checker for the HIP library, never part of
// it: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
// load it. Every function cites the reference file:line it follows.
//
// What it restates
//   * block grid and neighbour queries (Cubism Tree(info).Exists /
//     CheckCoarser / CheckFiner, Znei with periodic wrap, Zchild, Zparent),
//     as used at source/poisson/PoissonSolverExp.cpp:167-193,226-303;
//   * the sparse operator A, row by row, exactly as PoissonSolverExp::getMat
//     builds it (PoissonSolverExp.cpp:221-373), including makeFlux
//     (:157-219), interpolate (:112-155), D1/D2/Dmixed and the six face
//     indexers (PoissonSolverExp.h:63-478) and SpRowInfo's std::map column
//     merge (LocalSpMatDnVec.h:9-39);
//   * the block-Jacobi inverse P_inv by Cholesky-Crout + Gauss-Jordan
//     (PoissonSolverExp.cpp:17-107);
//   * the operator application with the bMeanConstraint row
//     (BiCGSTAB.cu:314-414) and getVec's zeroing (PoissonSolverExp.cpp:375-418);
//   * classical right-preconditioned BiCGSTAB (BiCGSTAB.cu:416-628);
//   * the CPU solver: pipelined BiCGSTAB (PoissonSolverAMR.cpp:13-358) with
//     the per-block CG preconditioner (PoissonSolverAMRKernels.cpp:135-332)
//     and the ComputeLHS bMean handling (PoissonSolverAMR.h:108-172).
//
// How each part is pinned (tests/test_oracle_*.py):
//   P_inv        closed-form DST eigendecomposition, P·A_loc = I, numpy inverse
//                fixture (tests/golden/pinv_fixture.json)
//   operator     periodic Fourier eigenpairs, Neumann cosine eigenpairs,
//                scipy brute-force 7-point matrix on a uniform grid, and on
//                AMR grids: zero row sums, exactness on quadratics away from
//                the domain walls, conservation 1^T A = 0, and a hand-derived
//                known answer for single coarse-fine rows
//   BiCGSTAB     converged solution vs scipy.sparse.linalg.spsolve, textbook
//                identities of the recurrences, bitwise determinism
//   pipelined    same converged solution as the classical solver
//   block CG     converges to the P_inv application (exact local inverse)
//
// Parity unpinned (nothing in the reference or in mathematics fixes it):
//   Cubism's block ordering (Z curve) — the caller's order is used as given;
//   the CPU "iterative" solver's AMR operator (Cubism BlockLab + flux
//   correction): the explicit GPU matrix is used for both solvers instead.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NX = 8, NY = 8, NZ = 8, NXYZ = 512;
enum { BC_FREESPACE = 0, BC_PERIODIC = 1, BC_WALL = 2 };  // Definitions.h:38

// ---------------------------------------------------------------------------
// Block grid: the minimal stand-in for Cubism's Grid/BlockInfo/Tree queries.
// ---------------------------------------------------------------------------
struct Blk {
  int level;
  int index[3];
};

struct Grid {
  int bpd[3];
  int levelMax;
  int bc[3];
  double h0;
  std::vector<Blk> blocks;
  std::unordered_map<uint64_t, int> leaf;  // (level,index) -> blockID

  static uint64_t key(int l, int i, int j, int k) {
    return (uint64_t(l) << 60) | (uint64_t(i) << 40) | (uint64_t(j) << 20) | uint64_t(k);
  }
  int nblocks() const { return (int)blocks.size(); }
  double h(int b) const { return h0 / double(1 << blocks[b].level); }
  int nbp(int l, int d) const { return bpd[d] << l; }  // blocks per dim at level l

  // Znei with periodic wrap (Cubism BlockInfo::Znei, used at
  // PoissonSolverExp.cpp:281-296). Wrapping is harmless for non-periodic
  // axes because boundary faces never query the neighbour (:361).
  void neighbour(int b, int face, int out[3]) const {
    const Blk& B = blocks[b];
    const int d = face / 2, s = (face % 2) ? 1 : -1;
    for (int a = 0; a < 3; ++a) out[a] = B.index[a];
    const int n = nbp(B.level, d);
    out[d] = ((out[d] + s) % n + n) % n;
  }
  int find(int l, const int I[3]) const {
    auto it = leaf.find(key(l, I[0], I[1], I[2]));
    return it == leaf.end() ? -1 : it->second;
  }
};

// ---------------------------------------------------------------------------
// Face indexers — PoissonSolverExp.h:63-371. face: 0 Xm,1 Xp,2 Ym,3 Yp,4 Zm,5 Zp
// ---------------------------------------------------------------------------
struct Indexer {
  const Grid& g;
  int face;
  long long This(int b, int ix, int iy, int iz) const {  // .h:69-70
    return (long long)b * NXYZ + (long long)(iz * NY * NX + iy * NX + ix);
  }
  static int ixf(int i) { return (i % (NX / 2)) * 2; }  // .h:102-104
  // Xmax/Xmin/... with offset (.h:86-97): the normal coordinate is pinned.
  long long faceCell(int b, int ix, int iy, int iz, int maxSide, int offset) const {
    const int d = face / 2;
    int c[3] = {ix, iy, iz};
    c[d] = maxSide ? (NX - 1 - offset) : offset;
    return This(b, c[0], c[1], c[2]);
  }
  // neiUnif (.h:180-181,207-208,252-253,279-280,324-325,351-352)
  long long neiUnif(int nb, int ix, int iy, int iz) const {
    return faceCell(nb, ix, iy, iz, (face % 2) == 0 /*min face -> max side*/, 0);
  }
  // neiInward (.h:183-184,210-211,...)
  long long neiInward(int b, int ix, int iy, int iz) const {
    int c[3] = {ix, iy, iz};
    c[face / 2] += (face % 2) ? -1 : 1;
    return This(b, c[0], c[1], c[2]);
  }
  // ix_c/iy_c/iz_c (.h:132-137, overridden along the normal at .h:186-187,...)
  int ic(int b, int i, int d) const {
    if (d == face / 2) return (face % 2) ? 0 : NX - 1;
    return (g.blocks[b].index[d] % 2 == 0) ? (i / 2) : (i / 2 + NX / 2);
  }
  // tangential directions: X faces ds1=y, ds2=z; Y faces ds1=z, ds2=x;
  // Z faces ds1=x, ds2=y (.h:158,230,302)
  int tdir(int dimShift) const {
    const int d = face / 2;
    if (d == 0) return dimShift == 1 ? 1 : 2;
    if (d == 1) return dimShift == 1 ? 2 : 0;
    return dimShift == 1 ? 0 : 1;
  }
  // neiFine1..4 (.h:189-196,...): order of (ds1,ds2) offsets (0,0),(1,0),(0,1),(1,1)
  long long neiFine(int nbf, int q, int ix, int iy, int iz, int offset) const {
    int c[3] = {ixf(ix), ixf(iy), ixf(iz)};
    if (q & 1) c[tdir(1)] += 1;
    if (q & 2) c[tdir(2)] += 1;
    return faceCell(nbf, c[0], c[1], c[2], (face % 2) == 0, offset);
  }
  double taylorSign(int ix, int iy, int iz, int dimShift) const {  // .h:164-165
    const int c[3] = {ix, iy, iz};
    return (c[tdir(dimShift)] % 2 == 0) ? -1. : 1.;
  }
  bool isBD(int ix, int iy, int iz, int dimShift) const {  // .h:166-167
    const int c[3] = {ix, iy, iz};
    const int t = c[tdir(dimShift)];
    return t == NX - 1 || t == NX / 2 - 1;
  }
  bool isFD(int ix, int iy, int iz, int dimShift) const {  // .h:168-169
    const int c[3] = {ix, iy, iz};
    const int t = c[tdir(dimShift)];
    return t == 0 || t == NX / 2;
  }
  long long Nei(int b, int ix, int iy, int iz, int d1, int d2) const {  // .h:170-171
    int c[3] = {ix, iy, iz};
    c[tdir(1)] += d1;
    c[tdir(2)] += d2;
    return This(b, c[0], c[1], c[2]);
  }
  // Zchild of the (absent) same-level neighbour adjacent to the face (.h:198-199,...)
  void Zchild(const int nI[3], int ix, int iy, int iz, int out[3]) const {
    const int half[3] = {ix >= NX / 2, iy >= NY / 2, iz >= NZ / 2};
    const int d = face / 2;
    for (int a = 0; a < 3; ++a) out[a] = 2 * nI[a] + half[a];
    out[d] = 2 * nI[d] + ((face % 2) ? 0 : 1);
  }
};

using Row = std::map<long long, double>;  // SpRowInfo::loc_colval_ (LocalSpMatDnVec.h:14)
typedef std::pair<long long, double> CV;

// D1, D2, Dmixed — PoissonSolverExp.h:384-478
static void D1(const Indexer& I, int b, int ix, int iy, int iz, int ds, CV out[3]) {
  const bool d1 = ds == 1;
  const int b2[2] = {d1 ? -2 : 0, d1 ? 0 : -2}, b1[2] = {d1 ? -1 : 0, d1 ? 0 : -1};
  const int f1[2] = {d1 ? 1 : 0, d1 ? 0 : 1}, f2[2] = {d1 ? 2 : 0, d1 ? 0 : 2};
  if (I.isBD(ix, iy, iz, ds)) {
    out[0] = {I.Nei(b, ix, iy, iz, b2[0], b2[1]), 1. / 8.};
    out[1] = {I.Nei(b, ix, iy, iz, b1[0], b1[1]), -1. / 2.};
    out[2] = {I.This(b, ix, iy, iz), 3. / 8.};
  } else if (I.isFD(ix, iy, iz, ds)) {
    out[0] = {I.Nei(b, ix, iy, iz, f2[0], f2[1]), -1. / 8.};
    out[1] = {I.Nei(b, ix, iy, iz, f1[0], f1[1]), 1. / 2.};
    out[2] = {I.This(b, ix, iy, iz), -3. / 8.};
  } else {
    out[0] = {I.Nei(b, ix, iy, iz, b1[0], b1[1]), -1. / 8.};
    out[1] = {I.Nei(b, ix, iy, iz, f1[0], f1[1]), 1. / 8.};
    out[2] = {I.This(b, ix, iy, iz), 0.};
  }
}
static void D2(const Indexer& I, int b, int ix, int iy, int iz, int ds, CV out[3]) {
  const bool d1 = ds == 1;
  const int b2[2] = {d1 ? -2 : 0, d1 ? 0 : -2}, b1[2] = {d1 ? -1 : 0, d1 ? 0 : -1};
  const int f1[2] = {d1 ? 1 : 0, d1 ? 0 : 1}, f2[2] = {d1 ? 2 : 0, d1 ? 0 : 2};
  if (I.isBD(ix, iy, iz, ds)) {
    out[0] = {I.Nei(b, ix, iy, iz, b2[0], b2[1]), 1. / 32.};
    out[1] = {I.Nei(b, ix, iy, iz, b1[0], b1[1]), -1. / 16.};
    out[2] = {I.This(b, ix, iy, iz), 1. / 32.};
  } else if (I.isFD(ix, iy, iz, ds)) {
    out[0] = {I.Nei(b, ix, iy, iz, f2[0], f2[1]), 1. / 32.};
    out[1] = {I.Nei(b, ix, iy, iz, f1[0], f1[1]), -1. / 16.};
    out[2] = {I.This(b, ix, iy, iz), 1. / 32.};
  } else {
    out[0] = {I.Nei(b, ix, iy, iz, b1[0], b1[1]), 1. / 32.};
    out[1] = {I.Nei(b, ix, iy, iz, f1[0], f1[1]), 1. / 32.};
    out[2] = {I.This(b, ix, iy, iz), -1. / 16.};
  }
}
static void Dmixed(const Indexer& I, int b, int ix, int iy, int iz, CV o[4]) {
  auto N = [&](int a, int c) { return I.Nei(b, ix, iy, iz, a, c); };
  const bool bd1 = I.isBD(ix, iy, iz, 1), bd2 = I.isBD(ix, iy, iz, 2);
  const bool fd1 = I.isFD(ix, iy, iz, 1), fd2 = I.isFD(ix, iy, iz, 2);
  const double s = 1. / 16., t = 1. / 32., u = 1. / 64.;
  if (bd1 && bd2) {
    o[0] = {N(0, 0), s}; o[1] = {N(-1, -1), s}; o[2] = {N(-1, 0), -s}; o[3] = {N(0, -1), -s};
  } else if (fd1 && bd2) {
    o[0] = {N(1, 0), s}; o[1] = {N(0, -1), s}; o[2] = {N(1, -1), -s}; o[3] = {N(0, 0), -s};
  } else if (bd1 && fd2) {
    o[0] = {N(0, 1), s}; o[1] = {N(-1, 0), s}; o[2] = {N(-1, 1), -s}; o[3] = {N(0, 0), -s};
  } else if (fd1 && fd2) {
    o[0] = {N(1, 1), s}; o[1] = {N(0, 0), s}; o[2] = {N(1, 0), -s}; o[3] = {N(0, 1), -s};
  } else if (bd1) {
    o[0] = {N(0, 1), t}; o[1] = {N(-1, -1), t}; o[2] = {N(0, -1), -t}; o[3] = {N(-1, 1), -t};
  } else if (fd1) {
    o[0] = {N(1, 1), t}; o[1] = {N(0, -1), t}; o[2] = {N(1, -1), -t}; o[3] = {N(0, 1), -t};
  } else if (bd2) {
    o[0] = {N(1, 0), t}; o[1] = {N(-1, -1), t}; o[2] = {N(-1, 0), -t}; o[3] = {N(1, -1), -t};
  } else if (fd2) {
    o[0] = {N(1, 1), t}; o[1] = {N(-1, 0), t}; o[2] = {N(-1, 1), -t}; o[3] = {N(1, 0), -t};
  } else {
    o[0] = {N(1, 1), u}; o[1] = {N(-1, -1), u}; o[2] = {N(1, -1), -u}; o[3] = {N(-1, 1), -u};
  }
}

// interpolate — PoissonSolverExp.cpp:112-155, with the factor sign * h_fine
// of the reference passed as `scale` (scale = 1 gives the ghost value's
// weights themselves).
static void interp_scaled(const Indexer& I, int bc, int ixc, int iyc, int izc, long long fclose,
                          long long ffar, double scale, double sds1, double sds2, Row& row) {
  row[fclose] += scale * 2. / 3.;
  row[ffar] += -scale * 1. / 5.;
  const double tf = scale * 8. / 15.;
  row[I.This(bc, ixc, iyc, izc)] += tf;
  CV D[3];
  D1(I, bc, ixc, iyc, izc, 1, D);
  for (int i = 0; i < 3; i++) row[D[i].first] += sds1 * tf * D[i].second;
  D1(I, bc, ixc, iyc, izc, 2, D);
  for (int i = 0; i < 3; i++) row[D[i].first] += sds2 * tf * D[i].second;
  D2(I, bc, ixc, iyc, izc, 1, D);
  for (int i = 0; i < 3; i++) row[D[i].first] += tf * D[i].second;
  D2(I, bc, ixc, iyc, izc, 2, D);
  for (int i = 0; i < 3; i++) row[D[i].first] += tf * D[i].second;
  CV Dm[4];
  Dmixed(I, bc, ixc, iyc, izc, Dm);
  for (int i = 0; i < 4; i++) row[Dm[i].first] += sds1 * sds2 * tf * Dm[i].second;
}
static void interpolate(const Grid& g, const Indexer& I, int bc, int ixc, int iyc, int izc,
                        int bf, long long fclose, long long ffar, double sign, double sds1,
                        double sds2, Row& row) {
  interp_scaled(I, bc, ixc, iyc, izc, fclose, ffar, sign * g.h(bf), sds1, sds2, row);
}

// makeFlux — PoissonSolverExp.cpp:157-219
static void makeFlux(const Grid& g, int b, int ix, int iy, int iz, int face, Row& row) {
  const Indexer I{g, face};
  const long long sfc = I.This(b, ix, iy, iz);
  int nI[3];
  g.neighbour(b, face, nI);
  const int l = g.blocks[b].level;
  const int same = g.find(l, nI);
  if (same >= 0) {  // Exists (:167-176)
    const double h = g.h(same);
    row[I.neiUnif(same, ix, iy, iz)] += h;
    row[sfc] += -h;
    return;
  }
  int pI[3] = {nI[0] / 2, nI[1] / 2, nI[2] / 2};
  const int coarse = l > 0 ? g.find(l - 1, pI) : -1;
  if (coarse >= 0) {  // CheckCoarser (:177-190)
    const int ixc = I.ic(b, ix, 0), iyc = I.ic(b, iy, 1), izc = I.ic(b, iz, 2);
    const long long inward = I.neiInward(b, ix, iy, iz);
    const double s1 = I.taylorSign(ix, iy, iz, 1), s2 = I.taylorSign(ix, iy, iz, 2);
    const double h = g.h(b);
    interpolate(g, I, coarse, ixc, iyc, izc, b, sfc, inward, 1., s1, s2, row);
    row[sfc] += -h;
    return;
  }
  int cI[3];  // CheckFiner (:191-217)
  I.Zchild(nI, ix, iy, iz, cI);
  const int fine = g.find(l + 1, cI);
  if (fine < 0) throw std::runtime_error("Neighbour doesn't exist, isn't coarser, nor finer...");
  const double h = g.h(fine);
  const double sgn[4][2] = {{-1., -1.}, {1., -1.}, {-1., 1.}, {1., 1.}};
  for (int q = 0; q < 4; ++q) {
    const long long fc = I.neiFine(fine, q, ix, iy, iz, 0);
    const long long ff = I.neiFine(fine, q, ix, iy, iz, 1);
    row[fc] += h;
    interpolate(g, I, b, ix, iy, iz, fine, fc, ff, -1., sgn[q][0], sgn[q][1], row);
  }
}

// ---------------------------------------------------------------------------
// The caller around the solve: KernelPressureRHS (PressureProjection.cpp:
// 96-199) and KernelGradP with the velocity update (:244-323, 449-466),
// including Cubism's flux correction at coarse-fine faces.
//
// Lab (ghost) values one cell outside a block, per face kind:
//   same level      the neighbour's cell (periodic wrap included);
//   coarser         the fine ghost of the operator's interface interpolation
//                   2/3 f_close - 1/5 f_far + 8/15 T(f_coarse) (interpolate
//                   above, scale 1): exact for quadratics (Cubism's BlockLab
//                   interpolation is absent from the reference mount);
//   domain boundary velocity (BlockLabBC, Definitions.h:284-435): the normal
//                   component is mirrored, ghost = -u (wall and freespace);
//                   pressure (BlockLabNeumann3D): ghost = p (zero gradient);
//   finer           the coarse cell's face term is replaced by the sum of the
//                   four fine cells' face terms, each built with its own fine
//                   ghost (the flux correction: the coarse side carries no
//                   flux of its own, then receives the fine fluxes).
// A face term is s fac (f_nb + f_c) for side s = -1 / +1, so the interior
// formula fac (f(x+h) - f(x-h)) is the sum of a cell's two face terms.
// ---------------------------------------------------------------------------
enum FaceClass { FC_INSIDE, FC_SAME, FC_COARSE, FC_FINE, FC_BOUNDARY };

// Class of face `face` of cell (ix,iy,iz) of block b, as makeFlux branches
// (PoissonSolverExp.cpp:157-219; boundaries as getMat, :286-296, 361).
static int face_class(const Grid& g, int b, int ix, int iy, int iz, int face, int& nb) {
  const int c[3] = {ix, iy, iz};
  const int d = face / 2, s = (face % 2) ? 1 : -1;
  if (c[d] + s >= 0 && c[d] + s < NX) return FC_INSIDE;
  const Blk& B = g.blocks[b];
  const bool boundary = s < 0 ? B.index[d] == 0 : B.index[d] == g.nbp(B.level, d) - 1;
  if (boundary && g.bc[d] != BC_PERIODIC) return FC_BOUNDARY;
  int nI[3];
  g.neighbour(b, face, nI);
  nb = g.find(B.level, nI);
  if (nb >= 0) return FC_SAME;
  const int pI[3] = {nI[0] / 2, nI[1] / 2, nI[2] / 2};
  nb = B.level > 0 ? g.find(B.level - 1, pI) : -1;
  if (nb >= 0) return FC_COARSE;
  return FC_FINE;
}

// f at cell k of a field with `nc` components, component `comp`.
static inline double fat(const double* f, int nc, int comp, long long k) {
  return f[k * nc + comp];
}
static double row_apply(const Row& r, const double* f, int nc, int comp) {
  double v = 0.;
  for (const auto& kv : r) v += kv.second * fat(f, nc, comp, kv.first);
  return v;
}

// Lab value across a same-level or coarser face of cell (b, ix, iy, iz).
static double ghost(const Grid& g, int b, int ix, int iy, int iz, int face, int cls, int nb,
                    const double* f, int nc, int comp) {
  const Indexer I{g, face};
  if (cls == FC_SAME) return fat(f, nc, comp, I.neiUnif(nb, ix, iy, iz));
  Row r;  // FC_COARSE: makeFlux's CheckCoarser branch with scale 1
  interp_scaled(I, nb, I.ic(b, ix, 0), I.ic(b, iy, 1), I.ic(b, iz, 2), I.This(b, ix, iy, iz),
                I.neiInward(b, ix, iy, iz), 1., I.taylorSign(ix, iy, iz, 1),
                I.taylorSign(ix, iy, iz, 2), r);
  return row_apply(r, f, nc, comp);
}

// Sum over the four fine cells behind a finer face of coarse cell (b, ix,
// iy, iz) of fac_f (G_q + f_q) - chi_q fac_f (Gs_q + fs_q), in the fine-cell
// order of makeFlux's CheckFiner branch (:191-217); fs / chi may be null.
static double fine_faces(const Grid& g, int b, int ix, int iy, int iz, int face, double fac_f,
                         const double* f, const double* fs, const double* chi, int nc,
                         int comp) {
  const Indexer I{g, face};
  int nI[3], cI[3];
  g.neighbour(b, face, nI);
  I.Zchild(nI, ix, iy, iz, cI);
  const int fine = g.find(g.blocks[b].level + 1, cI);
  if (fine < 0) throw std::runtime_error("finer face without a finer neighbour");
  const double sgn[4][2] = {{-1., -1.}, {1., -1.}, {-1., 1.}, {1., 1.}};
  double acc = 0.;
  for (int q = 0; q < 4; ++q) {
    const long long fc = I.neiFine(fine, q, ix, iy, iz, 0);
    const long long ff = I.neiFine(fine, q, ix, iy, iz, 1);
    Row r;
    interp_scaled(I, b, ix, iy, iz, fc, ff, 1., sgn[q][0], sgn[q][1], r);
    double t = (row_apply(r, f, nc, comp) + fat(f, nc, comp, fc));
    if (fs && chi) t -= chi[fc] * (row_apply(r, fs, nc, comp) + fat(fs, nc, comp, fc));
    acc += fac_f * t;
  }
  return acc;
}

// KernelPressureRHS (PressureProjection.cpp:109-199) with dt:
//   b = fac (uE - uW + vN - vS + wB - wF) - chi fac (the same of udef),
//   fac = h^2 / (2 dt), plus the flux corrections of finer faces.
// vel, udef: [n][3]; chi: [n]; udef and chi may be null (no obstacle).
static void pressure_rhs(const Grid& g, const double* vel, const double* udef,
                         const double* chi, double dt, double* out) {
  const int nb = g.nblocks();
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const double h = g.h(b), fac = 0.5 * h * h / dt;
    const double hf = 0.5 * h, fac_f = 0.5 * hf * hf / dt;
    const Indexer G{g, 0};
    for (int iz = 0; iz < NZ; ++iz)
      for (int iy = 0; iy < NY; ++iy)
        for (int ix = 0; ix < NX; ++ix) {
          const long long k = G.This(b, ix, iy, iz);
          double U[6], Us[6], corr = 0.;
          for (int face = 0; face < 6; ++face) {
            const int d = face / 2, s = (face % 2) ? 1 : -1;
            int nbk = -1;
            const int cls = face_class(g, b, ix, iy, iz, face, nbk);
            if (cls == FC_INSIDE) {
              const int c[3] = {ix + (d == 0 ? s : 0), iy + (d == 1 ? s : 0), iz + (d == 2 ? s : 0)};
              const long long kn = G.This(b, c[0], c[1], c[2]);
              U[face] = fat(vel, 3, d, kn);
              Us[face] = udef ? fat(udef, 3, d, kn) : 0.;
            } else if (cls == FC_BOUNDARY || cls == FC_FINE) {
              U[face] = -fat(vel, 3, d, k);
              Us[face] = udef ? -fat(udef, 3, d, k) : 0.;
              if (cls == FC_FINE)
                corr += s * fine_faces(g, b, ix, iy, iz, face, fac_f, vel, udef, chi, 3, d);
            } else {
              U[face] = ghost(g, b, ix, iy, iz, face, cls, nbk, vel, 3, d);
              Us[face] = udef ? ghost(g, b, ix, iy, iz, face, cls, nbk, udef, 3, d) : 0.;
            }
          }
          double v = fac * (U[1] - U[0] + U[3] - U[2] + U[5] - U[4]);
          if (udef && chi) v += -chi[k] * fac * (Us[1] - Us[0] + Us[3] - Us[2] + Us[5] - Us[4]);
          out[k] = v + corr;
        }
  }
}

// KernelGradP (PressureProjection.cpp:244-323) and the velocity update
// (:449-466): o = -dt h^2 / 2 (p(x+h) - p(x-h)) per component plus the flux
// corrections of finer faces; vel += o / h^3. vel: [n][3].
static void grad_p_update(const Grid& g, const double* p, double dt, double* vel) {
  const int nb = g.nblocks();
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const double h = g.h(b), fac = -0.5 * dt * h * h;
    const double hf = 0.5 * h, fac_f = -0.5 * dt * hf * hf;
    const double inv = 1.0 / (h * h * h);
    const Indexer G{g, 0};
    for (int iz = 0; iz < NZ; ++iz)
      for (int iy = 0; iy < NY; ++iy)
        for (int ix = 0; ix < NX; ++ix) {
          const long long k = G.This(b, ix, iy, iz);
          double P[6], corr[3] = {0., 0., 0.};
          for (int face = 0; face < 6; ++face) {
            const int d = face / 2, s = (face % 2) ? 1 : -1;
            int nbk = -1;
            const int cls = face_class(g, b, ix, iy, iz, face, nbk);
            if (cls == FC_INSIDE) {
              const int c[3] = {ix + (d == 0 ? s : 0), iy + (d == 1 ? s : 0), iz + (d == 2 ? s : 0)};
              P[face] = p[G.This(b, c[0], c[1], c[2])];
            } else if (cls == FC_BOUNDARY) {
              P[face] = p[k];
            } else if (cls == FC_FINE) {
              P[face] = -p[k];
              corr[d] += s * fine_faces(g, b, ix, iy, iz, face, fac_f, p, nullptr, nullptr, 1, 0);
            } else {
              P[face] = ghost(g, b, ix, iy, iz, face, cls, nbk, p, 1, 0);
            }
          }
          for (int d = 0; d < 3; ++d) {
            const double o = fac * (P[2 * d + 1] - P[2 * d]) + corr[d];
            vel[k * 3 + d] += inv * o;
          }
        }
  }
}

struct Matrix {
  int n = 0;
  std::vector<int64_t> rowptr;
  std::vector<int64_t> col;
  std::vector<double> val;
  int bmeanRow = -1;
};

// getMat — PoissonSolverExp.cpp:221-373 (single rank: Nblocks_xcumsum = 0)
static void assemble(const Grid& g, int bMean, Matrix& M) {
  const int nb = g.nblocks();
  M.n = nb * NXYZ;
  M.rowptr.assign(1, 0);
  M.col.clear();
  M.val.clear();
  M.col.reserve(size_t(M.n) * 7);
  M.val.reserve(size_t(M.n) * 7);
  M.bmeanRow = -1;
  auto push = [&](double v, long long c) {
    M.col.push_back(c);
    M.val.push_back(v);
  };
  for (int b = 0; b < nb; ++b) {
    const Blk& B = g.blocks[b];
    const int aux = 1 << B.level;
    bool isBoundary[6];
    for (int d = 0; d < 3; ++d) {
      isBoundary[2 * d] = B.index[d] == 0;
      isBoundary[2 * d + 1] = B.index[d] == g.bpd[d] * aux - 1;
    }
    const bool isPeriodic[3] = {g.bc[0] == BC_PERIODIC, g.bc[1] == BC_PERIODIC,
                                g.bc[2] == BC_PERIODIC};
    const bool origin = B.index[0] == 0 && B.index[1] == 0 && B.index[2] == 0;
    if (bMean == 1 && origin) M.bmeanRow = b * NXYZ;  // :299-303
    const Indexer G{g, 0};
    for (int iz = 0; iz < NZ; iz++)
      for (int iy = 0; iy < NY; iy++)
        for (int ix = 0; ix < NX; ix++) {
          const long long sfc = G.This(b, ix, iy, iz);
          const double h = g.h(b);
          if (bMean == 3 && origin && ix == 0 && iy == 0 && iz == 0) {  // :311-318
            push(h, sfc);
          } else if (ix > 0 && ix < NX - 1 && iy > 0 && iy < NY - 1 && iz > 0 && iz < NZ - 1) {
            push(h, G.This(b, ix, iy, iz - 1));  // :319-330
            push(h, G.This(b, ix, iy - 1, iz));
            push(h, G.This(b, ix - 1, iy, iz));
            push(-6. * h, sfc);
            push(h, G.This(b, ix + 1, iy, iz));
            push(h, G.This(b, ix, iy + 1, iz));
            push(h, G.This(b, ix, iy, iz + 1));
          } else {
            const bool valid[6] = {ix > 0, ix < NX - 1, iy > 0, iy < NY - 1, iz > 0, iz < NZ - 1};
            const long long idx[6] = {G.This(b, ix - 1, iy, iz), G.This(b, ix + 1, iy, iz),
                                      G.This(b, ix, iy - 1, iz), G.This(b, ix, iy + 1, iz),
                                      G.This(b, ix, iy, iz - 1), G.This(b, ix, iy, iz + 1)};
            Row row;
            for (int j = 0; j < 6; j++) {
              if (valid[j]) {
                row[idx[j]] += h;
                row[sfc] += -h;
              } else if (!isBoundary[j] || (isBoundary[j] && isPeriodic[j / 2])) {
                makeFlux(g, b, ix, iy, iz, j, row);
              }
            }
            for (auto& kv : row) push(kv.second, kv.first);
          }
          M.rowptr.push_back((int64_t)M.col.size());
        }
  }
}

// y = A x (local COO SpMV, BiCGSTAB.cu:339-349) followed by the bMean
// handling of hd_cusparseSpMV (BiCGSTAB.cu:392-413).
static void spmv(const Grid& g, const Matrix& M, int bMean, const double* x, double* y) {
  const int n = M.n;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int64_t k = M.rowptr[i]; k < M.rowptr[i + 1]; ++k) s += M.val[k] * x[M.col[k]];
    y[i] = s;
  }
  if (bMean == 1 || bMean == 2) {
    double red = 0;
    for (int b = 0; b < g.nblocks(); ++b) {
      const double h = g.h(b), h3 = h * h * h;
      for (int c = 0; c < NXYZ; ++c) red += h3 * x[(size_t)b * NXYZ + c];
    }
    if (bMean == 1 && M.bmeanRow >= 0)
      y[M.bmeanRow] = red;
    else if (bMean == 2)
      for (int b = 0; b < g.nblocks(); ++b) {
        const double h = g.h(b), h3 = h * h * h;
        for (int c = 0; c < NXYZ; ++c) y[(size_t)b * NXYZ + c] += red * h3;
      }
  }
}

// getA_local + Cholesky-Crout + Gauss-Jordan — PoissonSolverExp.cpp:17-107.
// `diag` is the diagonal of -A_loc: 6 for the pressure equation; 6 + h^2 /
// (nu dt) gives the per-level block inverse of the implicit diffusion
// operator (DiffusionSolverAMRKernels.cpp:325 coefficient, sign flipped the
// same way), whose reference preconditioner iterates CG towards it.
static double getA_local(int i, int j, double diag = 6.) {
  if (i == j) return diag;
  const int ix = i % NX, iy = (i / NX) % NY, iz = i / (NX * NY);
  const int jx = j % NX, jy = (j / NX) % NY, jz = j / (NX * NY);
  if (std::abs(ix - jx) + std::abs(iy - jy) + std::abs(iz - jz) == 1) return -1.;
  return 0.;
}
static void build_pinv(double* P, double diag = 6.) {
  const int n = NXYZ;
  std::vector<std::vector<double>> L(n), Li(n);
  for (int i = 0; i < n; i++) {
    L[i].assign(i + 1, 0.);
    Li[i].assign(i + 1, 0.);
    Li[i][i] = 1.;
  }
  for (int i = 0; i < n; i++) {
    double s1 = 0;
    for (int k = 0; k <= i - 1; k++) s1 += L[i][k] * L[i][k];
    L[i][i] = std::sqrt(getA_local(i, i, diag) - s1);
    for (int j = i + 1; j < n; j++) {
      double s2 = 0;
      for (int k = 0; k <= i - 1; k++) s2 += L[i][k] * L[j][k];
      L[j][i] = (getA_local(j, i, diag) - s2) / L[i][i];
    }
  }
  for (int br = 0; br < n; br++) {
    const double bsf = 1. / L[br][br];
    for (int c = 0; c <= br; c++) Li[br][c] *= bsf;
    for (int wr = br + 1; wr < n; wr++) {
      const double wsf = L[wr][br];
      for (int c = 0; c <= br; c++) Li[wr][c] -= wsf * Li[br][c];
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double aux = 0.;
      for (int k = std::max(i, j); k < n; k++) aux += Li[k][i] * Li[k][j];
      P[i * n + j] = -aux;
    }
}

struct Oracle {
  Grid g;
  Matrix M;
  int bMean = 1;
  std::vector<double> P;  // 512x512 row-major
  std::vector<double> h3, invh;
};

// z_b = invh_b * (P_inv p_b): cublasDgemm(OP_T,OP_N,...) + blockDscal
// (BiCGSTAB.cu:526-528).
static void precond(const Oracle& O, const double* p, double* z) {
  const int nb = O.g.nblocks();
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const double* pb = p + (size_t)b * NXYZ;
    double* zb = z + (size_t)b * NXYZ;
    for (int i = 0; i < NXYZ; ++i) {
      const double* Pi = O.P.data() + (size_t)i * NXYZ;
      double s = 0;
      for (int j = 0; j < NXYZ; ++j) s += Pi[j] * pb[j];
      zb[i] = s;
    }
    for (int i = 0; i < NXYZ; ++i) zb[i] = O.invh[b] * zb[i];
  }
}

static double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
static double nrm2(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }
static void axpy(double a, const std::vector<double>& x, std::vector<double>& y) {
  for (size_t i = 0; i < x.size(); ++i) y[i] = a * x[i] + y[i];
}

// getVec — PoissonSolverExp.cpp:375-418: b and x copied, zeroed at the bMean cell.
static void get_vec(const Oracle& O, const double* rhs, const double* x0, std::vector<double>& b,
                    std::vector<double>& x) {
  const int n = O.M.n;
  b.assign(rhs, rhs + n);
  if (x0)
    x.assign(x0, x0 + n);
  else
    x.assign(n, 0.);
  if (O.bMean == 1 || O.bMean == 3)
    for (int bb = 0; bb < O.g.nblocks(); ++bb) {
      const Blk& B = O.g.blocks[bb];
      if (B.index[0] == 0 && B.index[1] == 0 && B.index[2] == 0) {
        b[(size_t)bb * NXYZ] = 0.;
        x[(size_t)bb * NXYZ] = 0.;
      }
    }
}

struct BiRes {
  int status, iters, restarts;
  double err_init, err_opt;
};

// BiCGSTABSolver::main — BiCGSTAB.cu:416-628, over an operator `apply`
// (y = A x) and a preconditioner `prec` (z = M^-1 p); b and x hold getVec's
// rhs and initial guess.
template <class Apply, class Prec>
static BiRes bicgstab_core(int n, const std::vector<double>& b, std::vector<double> x,
                           double* xout, Apply&& apply, Prec&& prec, double max_error,
                           double max_rel_error, int max_restarts, int max_iter, double* hist) {
  std::vector<double> r, rhat, p(n, 0.), nu(n, 0.), t(n, 0.), z(n, 0.), x_opt;
  r = b;
  struct {
    double alpha, beta, omega, eps, rho_prev, rho_curr, buff_1, buff_2;
  } c = {1., 1., 1., 1e-21, 1., 1., 0., 0.};  // :431
  double error, error_init, error_opt;
  bool converged = false;
  int restarts = 0;
  z = x;  // :435
  apply(z.data(), nu.data());
  axpy(-1., nu, r);  // :437
  c.buff_1 = nrm2(nu);
  c.buff_2 = nrm2(r);
  c.buff_1 *= c.buff_1;
  c.buff_2 *= c.buff_2;
  c.buff_1 = std::sqrt(c.buff_1);
  c.buff_2 = std::sqrt(c.buff_2);
  error = c.buff_2;
  error_init = error;
  error_opt = error;
  x_opt = x;
  rhat = r;
  std::fill(nu.begin(), nu.end(), 0.);
  std::fill(p.begin(), p.end(), 0.);
  int k = 0, iters = 0;
  for (k = 0; k < max_iter; k++) {
    c.rho_curr = dot(rhat, r);  // :478
    c.buff_1 = nrm2(r);
    c.buff_2 = nrm2(rhat);
    c.buff_1 *= c.buff_1;
    c.buff_2 *= c.buff_2;
    const bool serious_breakdown = c.rho_curr * c.rho_curr < 1e-16 * c.buff_1 * c.buff_2;
    c.beta = (c.rho_curr / (c.rho_prev + c.eps)) * (c.alpha / (c.omega + c.eps));  // :492
    if (serious_breakdown && max_restarts > 0) {
      restarts++;
      if (restarts >= max_restarts) break;
      rhat = r;
      c.rho_curr = nrm2(rhat);
      c.rho_curr *= c.rho_curr;
      std::fill(nu.begin(), nu.end(), 0.);
      std::fill(p.begin(), p.end(), 0.);
      c.rho_prev = 1.;  // breakdown_update (:263-269)
      c.alpha = 1.;
      c.omega = 1.;
      c.beta = (c.rho_curr / (c.rho_prev + c.eps)) * (c.alpha / (c.omega + c.eps));
    }
    iters = k + 1;
    c.buff_1 = -c.omega;  // :518-522
    axpy(c.buff_1, nu, p);
    for (int i = 0; i < n; ++i) p[i] = c.beta * p[i];
    axpy(1., r, p);
    prec(p.data(), z.data());                   // :526-528
    apply(z.data(), nu.data());      // :532
    c.buff_1 = dot(rhat, nu);                          // :535
    c.alpha = c.rho_curr / (c.buff_1 + c.eps);         // :540
    axpy(c.alpha, z, x);                               // :544
    axpy(-c.alpha, nu, r);                             // :547-549
    prec(r.data(), z.data());                    // :553-555
    apply(z.data(), t.data());       // :559
    c.buff_1 = dot(t, r);                              // :562
    c.buff_2 = nrm2(t);
    c.buff_2 *= c.buff_2;
    c.omega = c.buff_1 / (c.buff_2 + c.eps);           // :569
    axpy(c.omega, z, x);                               // :573
    axpy(-c.omega, t, r);                              // :576-578
    error = nrm2(r);                                   // :581-588
    if (hist) hist[k] = error;
    if (error < error_opt) {                           // :590-603
      error_opt = error;
      x_opt = x;
      if ((error <= max_error) || (error / error_init <= max_rel_error)) {
        converged = true;
        break;
      }
    }
    c.rho_prev = c.rho_curr;  // :607
  }
  std::copy(x_opt.begin(), x_opt.end(), xout);
  BiRes res;
  res.status = converged ? 0 : (restarts >= max_restarts && max_restarts > 0 ? 2 : 1);
  res.iters = iters;
  res.restarts = restarts;
  res.err_init = error_init;
  res.err_opt = error_opt;
  return res;
}

static BiRes bicgstab(const Oracle& O, const double* rhs, const double* x0, double* xout,
                      double max_error, double max_rel_error, int max_restarts, int max_iter,
                      double* hist) {
  std::vector<double> b, x;
  get_vec(O, rhs, x0, b, x);
  return bicgstab_core(
      O.M.n, b, std::move(x), xout, [&](const double* a, double* y) { spmv(O.g, O.M, O.bMean, a, y); },
      [&](const double* a, double* y) { precond(O, a, y); }, max_error, max_rel_error,
      max_restarts, max_iter, hist);
}

// ---------------------------------------------------------------------------
// CPU solver: per-block CG preconditioner (PoissonSolverAMRKernels.cpp:135-332)
// ---------------------------------------------------------------------------
// `sigma` = 1 / (nu dt) turns it into the diffusion solver's preconditioner
// (DiffusionSolverAMRKernels.cpp:149-213): the same CG with the diagonal
// coefficient -6 - h^2 sigma (:206) in place of -6; sigma = 0 is the pressure
// solver's.
static void block_cg(const Oracle& O, const double* in, double* out, double sigma = 0.) {
  const int nb = O.g.nblocks();
  constexpr double kDivEpsilon = 1e-55;           // :76
  constexpr double kNormRel = 1e-7, kNormAbs = 1e-16;  // :77-78
  constexpr double kSqrRel = kNormRel * kNormRel, kSqrAbs = kNormAbs * kNormAbs;
  constexpr double Ncells = NXYZ;
#pragma omp parallel for schedule(static)
  for (int bb = 0; bb < nb; ++bb) {
    double r[NZ][NY][NX], Ax[NZ][NY][NX], p[NZ + 2][NY + 2][NX + 2];
    double* blk = out + (size_t)bb * NXYZ;
    const double* src = in + (size_t)bb * NXYZ;
    std::memset(p, 0, sizeof(p));
    const double invh = 1 / O.g.h(bb);
    const double coef = -6.0 - O.g.h(bb) * O.g.h(bb) * sigma;
    double rr = 0;
    for (int iz = 0; iz < NZ; ++iz)  // :305-315
      for (int iy = 0; iy < NY; ++iy)
        for (int ix = 0; ix < NX; ++ix) {
          const int j = iz * 64 + iy * 8 + ix;
          r[iz][iy][ix] = invh * src[j];
          rr += r[iz][iy][ix] * r[iz][iy][ix];
          p[iz + 1][iy + 1][ix + 1] = r[iz][iy][ix];
          blk[j] = 0;
        }
    const double sqrNorm0 = 1. / (Ncells * Ncells) * rr;  // :317
    if (sqrNorm0 < 1e-32) continue;                        // :319-320
    for (int k = 0; k < 100; ++k) {                        // :325-330
      double a2 = 0;                                       // :196-232
      for (int iz = 0; iz < NZ; ++iz)
        for (int iy = 0; iy < NY; ++iy)
          for (int ix = 0; ix < NX; ++ix) {
            const double v = p[iz + 1][iy + 1][ix] + p[iz + 1][iy + 1][ix + 2] +
                             coef * p[iz + 1][iy + 1][ix + 1] + p[iz + 1][iy][ix + 1] +
                             p[iz + 1][iy + 2][ix + 1] + p[iz][iy + 1][ix + 1] +
                             p[iz + 2][iy + 1][ix + 1];
            Ax[iz][iy][ix] = v;
            a2 += p[iz + 1][iy + 1][ix + 1] * v;
          }
      const double a = rr / (a2 + kDivEpsilon);  // :233
      double sqrSum = 0;
      for (int iz = 0; iz < NZ; ++iz)  // :239-247
        for (int iy = 0; iy < NY; ++iy)
          for (int ix = 0; ix < NX; ++ix) {
            blk[iz * 64 + iy * 8 + ix] += a * p[iz + 1][iy + 1][ix + 1];
            r[iz][iy][ix] -= a * Ax[iz][iy][ix];
            sqrSum += r[iz][iy][ix] * r[iz][iy][ix];
          }
      const double beta = sqrSum / (rr + kDivEpsilon);  // :249
      const double sqrNorm = 1. / (Ncells * Ncells) * sqrSum;
      if (sqrNorm < kSqrRel * sqrNorm0 || sqrNorm < kSqrAbs) break;  // :252-253
      for (int iz = 0; iz < NZ; ++iz)
        for (int iy = 0; iy < NY; ++iy)
          for (int ix = 0; ix < NX; ++ix)
            p[iz + 1][iy + 1][ix + 1] = r[iz][iy][ix] + beta * p[iz + 1][iy + 1][ix + 1];
      rr = sqrSum;
      if (rr <= 0) break;
    }
  }
}

// ComputeLHS (PoissonSolverAMR.h:108-172): A x, then row (0,0,0) of block
// index (0,0,0) := sum h^3 x for bMean==1, += (sum h^3 x) h^3 for bMean==2,
// := x(0,0,0) for bMean>2.
static void lhs_cpu(const Oracle& O, const double* x, double* y) {
  const Matrix& M = O.M;
  const int n = M.n;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int64_t k = M.rowptr[i]; k < M.rowptr[i + 1]; ++k) s += M.val[k] * x[M.col[k]];
    y[i] = s;
  }
  const int bm = O.bMean;
  if (bm == 0) return;
  int origin = -1;
  for (int b = 0; b < O.g.nblocks(); ++b) {
    const Blk& B = O.g.blocks[b];
    if (B.index[0] == 0 && B.index[1] == 0 && B.index[2] == 0) origin = b;
  }
  if (bm <= 2) {
    double avg = 0;
#pragma omp parallel for reduction(+ : avg) schedule(static)
    for (int b = 0; b < O.g.nblocks(); ++b) {
      double s = 0;
      for (int c = 0; c < NXYZ; ++c) s += x[(size_t)b * NXYZ + c] * O.h3[b];
      avg += s;
    }
    if (bm == 1 && origin >= 0)
      y[(size_t)origin * NXYZ] = avg;
    else if (bm == 2)
      for (int b = 0; b < O.g.nblocks(); ++b)
        for (int c = 0; c < NXYZ; ++c) y[(size_t)b * NXYZ + c] += avg * O.h3[b];
  } else if (origin >= 0) {
    y[(size_t)origin * NXYZ] = x[(size_t)origin * NXYZ];
  }
}

// PoissonSolverAMR::solve — PoissonSolverAMR.cpp:13-358 (Cools-Vanroose
// Alg. 11), over an operator `lhs` and a preconditioner `prec`; b holds the
// rhs (after the bMean zeroing) and x the initial guess. The implicit
// diffusion solver DiffusionSolver::solve (AdvectionDiffusionImplicit.h:
// 219-525) is the same algorithm with unlimited restarts (its restart cap is
// commented out, :487).
template <class Lhs, class Prec>
static BiRes pipelined_core(size_t N, const std::vector<double>& b, std::vector<double> x,
                            double* xout, Lhs&& lhs, Prec&& prec, double max_error,
                            double max_rel_error, int max_iter, int max_restarts) {
  const double eps = 1e-100;
  std::vector<double> phat(N), rhat(N), shat(N), what(N), zhat(N), qhat(N), s(N), w(N), z(N),
      t(N), v(N), q(N), r(b), y(N), r0(N), x_opt(N);
  bool useXopt = false;
  int restarts = 0;
  double min_norm = 1e50, norm_1 = 0, norm_2 = 0;
  lhs(x.data(), r0.data());
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    r0[i] = r[i] - r0[i];
    r[i] = r0[i];
  }
  prec(r0.data(), rhat.data());
  lhs(rhat.data(), w.data());
  prec(w.data(), what.data());
  lhs(what.data(), t.data());
  double alpha = 0, norm = 0, beta = 0, omega = 0, r0r_prev;
  {
    double temp0 = 0, temp1 = 0;
#pragma omp parallel for reduction(+ : temp0, temp1, norm)
    for (size_t j = 0; j < N; j++) {
      temp0 += r0[j] * r0[j];
      temp1 += r0[j] * w[j];
      norm += r0[j] * r0[j];
    }
    alpha = temp0 / (temp1 + eps);
    r0r_prev = temp0;
    norm = std::sqrt(norm);
  }
  const double init_norm = norm;
  int k, iters = 0;
  bool converged = false;
  for (k = 0; k < max_iter; k++) {
    iters = k + 1;
    double qy = 0, yy = 0;
    if (k % 50 != 0) {
#pragma omp parallel for reduction(+ : qy, yy)
      for (size_t j = 0; j < N; j++) {
        phat[j] = rhat[j] + beta * (phat[j] - omega * shat[j]);
        s[j] = w[j] + beta * (s[j] - omega * z[j]);
        shat[j] = what[j] + beta * (shat[j] - omega * zhat[j]);
        z[j] = t[j] + beta * (z[j] - omega * v[j]);
        q[j] = r[j] - alpha * s[j];
        qhat[j] = rhat[j] - alpha * shat[j];
        y[j] = w[j] - alpha * z[j];
        qy += q[j] * y[j];
        yy += y[j] * y[j];
      }
    } else {
#pragma omp parallel for
      for (size_t j = 0; j < N; j++) phat[j] = rhat[j] + beta * (phat[j] - omega * shat[j]);
      lhs(phat.data(), s.data());
      prec(s.data(), shat.data());
      lhs(shat.data(), z.data());
#pragma omp parallel for reduction(+ : qy, yy)
      for (size_t j = 0; j < N; j++) {
        q[j] = r[j] - alpha * s[j];
        qhat[j] = rhat[j] - alpha * shat[j];
        y[j] = w[j] - alpha * z[j];
        qy += q[j] * y[j];
        yy += y[j] * y[j];
      }
    }
    prec(z.data(), zhat.data());
    lhs(zhat.data(), v.data());
    omega = qy / (yy + eps);
    double r0r = 0, r0w = 0, r0s = 0, r0z = 0;
    norm = 0;
    norm_1 = 0;
    norm_2 = 0;
    if (k % 50 != 0) {
#pragma omp parallel for reduction(+ : r0r, r0w, r0s, r0z, norm_1, norm_2, norm)
      for (size_t j = 0; j < N; j++) {
        x[j] = x[j] + alpha * phat[j] + omega * qhat[j];
        r[j] = q[j] - omega * y[j];
        rhat[j] = qhat[j] - omega * (what[j] - alpha * zhat[j]);
        w[j] = y[j] - omega * (t[j] - alpha * v[j]);
        r0r += r0[j] * r[j];
        r0w += r0[j] * w[j];
        r0s += r0[j] * s[j];
        r0z += r0[j] * z[j];
        norm += r[j] * r[j];
        norm_1 += r[j] * r[j];
        norm_2 += r0[j] * r0[j];
      }
    } else {
#pragma omp parallel for
      for (size_t j = 0; j < N; j++) x[j] = x[j] + alpha * phat[j] + omega * qhat[j];
      lhs(x.data(), r.data());
#pragma omp parallel for
      for (size_t j = 0; j < N; j++) r[j] = b[j] - r[j];
      prec(r.data(), rhat.data());
      lhs(rhat.data(), w.data());
#pragma omp parallel for reduction(+ : r0r, r0w, r0s, r0z, norm_1, norm_2, norm)
      for (size_t j = 0; j < N; j++) {
        r0r += r0[j] * r[j];
        r0w += r0[j] * w[j];
        r0s += r0[j] * s[j];
        r0z += r0[j] * z[j];
        norm += r[j] * r[j];
        norm_1 += r[j] * r[j];
        norm_2 += r0[j] * r0[j];
      }
    }
    prec(w.data(), what.data());
    lhs(what.data(), t.data());
    norm = std::sqrt(norm);
    beta = alpha / (omega + eps) * r0r / (r0r_prev + eps);  // :275
    alpha = r0r / (r0w + beta * r0s - beta * omega * r0z);   // :278-281
    double alphat = 1.0 / (omega + eps) + r0w / (r0r + eps) - beta * omega * r0z / (r0r + eps);
    alphat = 1.0 / (alphat + eps);
    if (std::fabs(alphat) < 10 * std::fabs(alpha)) alpha = alphat;
    r0r_prev = r0r;
    const bool serious_breakdown = r0r * r0r < 1e-16 * norm_1 * norm_2;  // :286
    if (serious_breakdown && restarts < max_restarts) {
      restarts++;
      for (size_t i = 0; i < N; i++) r0[i] = r[i];
      prec(r0.data(), rhat.data());
      lhs(rhat.data(), w.data());
      double temp0 = 0, temp1 = 0;
      for (size_t j = 0; j < N; j++) {
        temp0 += r0[j] * r0[j];
        temp1 += r0[j] * w[j];
      }
      prec(w.data(), what.data());
      lhs(what.data(), t.data());
      alpha = temp0 / (temp1 + eps);
      r0r_prev = temp0;
      beta = 0.0;
      omega = 0.0;
    }
    if (norm < min_norm) {  // :324-331
      useXopt = true;
      min_norm = norm;
      x_opt = x;
    }
    if (norm < max_error || norm / (init_norm + eps) < max_rel_error) {  // :332-337
      converged = true;
      break;
    }
  }
  const std::vector<double>& sol = useXopt ? x_opt : x;
  std::copy(sol.begin(), sol.end(), xout);
  BiRes res;
  res.status = converged ? 0 : 1;
  res.iters = iters;
  res.restarts = restarts;
  res.err_init = init_norm;
  res.err_opt = min_norm;
  return res;
}


static BiRes pipelined(const Oracle& O, const double* rhs, const double* x0, double* xout,
                       double max_error, double max_rel_error, int max_iter) {
  const size_t N = O.M.n;
  int origin = -1;
  for (int bb = 0; bb < O.g.nblocks(); ++bb) {
    const Blk& B = O.g.blocks[bb];
    if (B.index[0] == 0 && B.index[1] == 0 && B.index[2] == 0) origin = bb;
  }
  std::vector<double> b(rhs, rhs + N), x(N, 0.);
  if (x0) std::copy(x0, x0 + N, x.begin());
  if ((O.bMean == 1 || O.bMean > 2) && origin >= 0) b[(size_t)origin * NXYZ] = 0;  // :68-70
  return pipelined_core(
      N, b, std::move(x), xout, [&](const double* a, double* y) { lhs_cpu(O, a, y); },
      [&](const double* a, double* y) { block_cg(O, a, y); }, max_error, max_rel_error,
      max_iter, 100);
}

// ---------------------------------------------------------------------------
// Implicit diffusion: the operator of KernelLHSDiffusion
// (AdvectionDiffusionImplicit.h:40-110), the per-block CG preconditioner
// (DiffusionSolverAMRKernels.cpp:149-213, block_cg above with sigma), the
// pipelined solver DiffusionSolver::solve (AdvectionDiffusionImplicit.h:
// 219-525) and the step around it (AdvectionDiffusionImplicit.cpp:354-419,
// from after the advection kernel). sigma = 1 / (nu dt) throughout.
//
// Lab values per face, for velocity component `comp` (BlockLabBC<..., comp>,
// Definitions.h:58-122 open faces, :287-330 walls, :413-458 dispatch):
//   inside / same level   the neighbouring cell (periodic wrap included);
//   coarser               the fine ghost of the interface interpolation, as
//                         for the pressure operator (Cubism's BlockLab
//                         interpolation is absent from the reference mount);
//   finer                 the coarse cell's face flux is replaced by the sum
//                         of the four fine cells' fluxes h_f (f_q - G_q)
//                         (Cubism's FluxCorrection, the same flux the
//                         pressure operator's coarse rows carry);
//   domain boundary       wall: ghost = -f (u = v = w = 0); freespace: ghost
//                         = -f for the component normal to the face, +f
//                         for the tangential ones.
// ---------------------------------------------------------------------------
static double bc_ghost_sign(const Grid& g, int face, int comp) {
  const int d = face / 2;
  if (g.bc[d] == BC_WALL) return -1.;
  return d == comp ? -1. : 1.;
}

// Sum over the four fine cells behind a finer face of coarse cell (b, ix,
// iy, iz) of (f_q - G_q), G_q the fine ghost, in makeFlux's fine-cell order.
static double fine_flux_sum(const Grid& g, int b, int ix, int iy, int iz, int face,
                            const double* f, int nc, int comp) {
  const Indexer I{g, face};
  int nI[3], cI[3];
  g.neighbour(b, face, nI);
  I.Zchild(nI, ix, iy, iz, cI);
  const int fine = g.find(g.blocks[b].level + 1, cI);
  if (fine < 0) throw std::runtime_error("finer face without a finer neighbour");
  const double sgn[4][2] = {{-1., -1.}, {1., -1.}, {-1., 1.}, {1., 1.}};
  double acc = 0.;
  for (int q = 0; q < 4; ++q) {
    const long long fc = I.neiFine(fine, q, ix, iy, iz, 0);
    const long long ff = I.neiFine(fine, q, ix, iy, iz, 1);
    Row r;
    interp_scaled(I, b, ix, iy, iz, fc, ff, 1., sgn[q][0], sgn[q][1], r);
    acc += fat(f, nc, comp, fc) - row_apply(r, f, nc, comp);
  }
  return acc;
}

// out = h (sum_faces (lab - f)) - h^3 sigma f for component `fcomp` of an
// `nc`-component field f, with the boundary ghosts of velocity component
// `bccomp`. sigma = 0 is KernelDiffusionRHS's h lap(u) (facD = h,
// AdvectionDiffusionImplicit.cpp:33-44 and its flux correction :46-123).
static void diffusion_lhs(const Grid& g, int bccomp, double sigma, const double* f, int nc,
                          int fcomp, double* out) {
  const int nb = g.nblocks();
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const double h = g.h(b), hf = 0.5 * h, coef = -h * h * h * sigma;
    const Indexer G{g, 0};
    for (int iz = 0; iz < NZ; ++iz)
      for (int iy = 0; iy < NY; ++iy)
        for (int ix = 0; ix < NX; ++ix) {
          const long long k = G.This(b, ix, iy, iz);
          const double fc = fat(f, nc, fcomp, k);
          double acc = 0.;
          for (int face = 0; face < 6; ++face) {
            const int d = face / 2, s = (face % 2) ? 1 : -1;
            int nbk = -1;
            const int cls = face_class(g, b, ix, iy, iz, face, nbk);
            if (cls == FC_INSIDE) {
              const int c[3] = {ix + (d == 0 ? s : 0), iy + (d == 1 ? s : 0), iz + (d == 2 ? s : 0)};
              acc += h * (fat(f, nc, fcomp, G.This(b, c[0], c[1], c[2])) - fc);
            } else if (cls == FC_BOUNDARY) {
              acc += h * (bc_ghost_sign(g, face, bccomp) * fc - fc);
            } else if (cls == FC_FINE) {
              acc += hf * fine_flux_sum(g, b, ix, iy, iz, face, f, nc, fcomp);
            } else {
              acc += h * (ghost(g, b, ix, iy, iz, face, cls, nbk, f, nc, fcomp) - fc);
            }
          }
          out[k] = acc + coef * fc;
        }
  }
}

struct DiffusionPinv {
  std::map<int, std::vector<double>> by_level;  // P_inv(6 + h_l^2 sigma) per level
};

// z_b = invh_b P_inv(level_b) p_b: the exact per-level block inverse the
// reference's CG preconditioner iterates towards.
static void diffusion_precond(const Oracle& O, double sigma, DiffusionPinv& D, const double* p,
                              double* z) {
  const int nb = O.g.nblocks();
  for (int b = 0; b < nb; ++b) {
    const int l = O.g.blocks[b].level;
    if (!D.by_level.count(l)) {
      std::vector<double> P((size_t)NXYZ * NXYZ);
      const double h = O.g.h(b);
      build_pinv(P.data(), 6. + h * h * sigma);
      D.by_level[l] = std::move(P);
    }
  }
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    const double* P = D.by_level.at(O.g.blocks[b].level).data();
    const double* pb = p + (size_t)b * NXYZ;
    double* zb = z + (size_t)b * NXYZ;
    for (int i = 0; i < NXYZ; ++i) {
      double acc = 0;
      for (int j = 0; j < NXYZ; ++j) acc += P[(size_t)i * NXYZ + j] * pb[j];
      zb[i] = O.invh[b] * acc;
    }
  }
}

// One implicit-diffusion step after the advection kernel
// (AdvectionDiffusionImplicit.cpp:354-419): vel_old = u^n, vel = u* (in) and
// u^{n+1} (out), both [n][3]. TMPV = h lap(u*) for all three components
// first; then per component: rhs = h^3 (-TMPV / h^3 + (u* - u^n) / (dt nu)),
// x0 = 0, the pipelined solve, u*_c += x.
static void diffusion_step(const Oracle& O, double nu, double dt, double tol_abs,
                           double tol_rel, int max_iter, const double* vel_old, double* vel,
                           int* iters, double* errs) {
  const size_t n = O.M.n;
  const double sigma = 1.0 / (nu * dt);
  std::vector<double> tmp(3 * n);
  for (int c = 0; c < 3; ++c) diffusion_lhs(O.g, c, 0.0, vel, 3, c, tmp.data() + c * n);
  for (int c = 0; c < 3; ++c) {
    std::vector<double> b(n), x(n);
    for (int bb = 0; bb < O.g.nblocks(); ++bb) {
      const double h3 = O.h3[bb], ih3 = 1.0 / h3;
      for (int j = 0; j < NXYZ; ++j) {
        const size_t k = (size_t)bb * NXYZ + j;
        const double t = -tmp[c * n + k] * ih3 + (vel[3 * k + c] - vel_old[3 * k + c]) / (dt * nu);
        b[k] = h3 * t;
      }
    }
    BiRes r = pipelined_core(
        n, b, std::vector<double>(n, 0.), x.data(),
        [&](const double* a, double* y) { diffusion_lhs(O.g, c, sigma, a, 1, 0, y); },
        [&](const double* a, double* y) { block_cg(O, a, y, sigma); }, tol_abs, tol_rel,
        max_iter, 1 << 30);
    for (size_t k = 0; k < n; ++k) vel[3 * k + c] += x[k];
    iters[c] = r.iters;
    errs[2 * c] = r.err_init;
    errs[2 * c + 1] = r.err_opt;
  }
}

thread_local std::string g_err;

}  // namespace

// ---------------------------------------------------------------------------
// C interface (ctypes) — test infrastructure only.
// ---------------------------------------------------------------------------
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// blocks: nblocks x 4 ints (level, ix, iy, iz) in blockID order.
void* orc_create(const int* bpd, int levelMax, const int* bc, double h0, int bMean, int nblocks,
                 const int* blocks) {
  try {
    Oracle* O = new Oracle;
    for (int d = 0; d < 3; ++d) {
      O->g.bpd[d] = bpd[d];
      O->g.bc[d] = bc[d];
    }
    O->g.levelMax = levelMax;
    O->g.h0 = h0;
    O->bMean = bMean;
    O->g.blocks.resize(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      Blk& B = O->g.blocks[b];
      B.level = blocks[4 * b];
      for (int d = 0; d < 3; ++d) B.index[d] = blocks[4 * b + 1 + d];
      O->g.leaf[Grid::key(B.level, B.index[0], B.index[1], B.index[2])] = b;
    }
    assemble(O->g, bMean, O->M);
    O->h3.resize(nblocks);
    O->invh.resize(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      const double h = O->g.h(b);
      O->h3[b] = h * h * h;  // getVec (PoissonSolverExp.cpp:395-396)
      O->invh[b] = 1. / h;
    }
    O->P.resize((size_t)NXYZ * NXYZ);
    build_pinv(O->P.data());
    return O;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_destroy(void* o) { delete (Oracle*)o; }

int64_t orc_nnz(void* o) { return (int64_t)((Oracle*)o)->M.col.size(); }
int orc_bmean_row(void* o) { return ((Oracle*)o)->M.bmeanRow; }

void orc_get_csr(void* o, int64_t* rowptr, int64_t* col, double* val) {
  const Matrix& M = ((Oracle*)o)->M;
  std::copy(M.rowptr.begin(), M.rowptr.end(), rowptr);
  std::copy(M.col.begin(), M.col.end(), col);
  std::copy(M.val.begin(), M.val.end(), val);
}

void orc_get_pinv(void* o, double* P) {
  const Oracle* O = (Oracle*)o;
  std::copy(O->P.begin(), O->P.end(), P);
}

// Standalone P_inv (no grid needed).
void orc_pinv(double* P) { build_pinv(P); }

// y = A x with the GPU solver's bMean semantics (hd_cusparseSpMV).
void orc_spmv(void* o, const double* x, double* y) {
  const Oracle* O = (Oracle*)o;
  spmv(O->g, O->M, O->bMean, x, y);
}
// y = A x, plain matrix (no bMean row handling).
void orc_spmv_raw(void* o, const double* x, double* y) {
  const Oracle* O = (Oracle*)o;
  spmv(O->g, O->M, 0, x, y);
}
// z = blockdiag(invh_b P_inv) p
void orc_precond(void* o, const double* p, double* z) { precond(*(Oracle*)o, p, z); }
// z = per-block CG approximation of the same operator (CPU solver).
void orc_block_cg(void* o, const double* p, double* z) { block_cg(*(Oracle*)o, p, z); }
// ComputeLHS with its bMean handling (CPU solver operator).
void orc_lhs_cpu(void* o, const double* x, double* y) { lhs_cpu(*(Oracle*)o, x, y); }

// Classical BiCGSTAB (BiCGSTAB.cu:416-628). out: iters, restarts; errs[2] = {err_init, err_opt}
int orc_bicgstab(void* o, const double* rhs, const double* x0, double* x, double tol_abs,
                 double tol_rel, int max_restarts, int max_iter, int* iters_restarts,
                 double* errs, double* hist) {
  BiRes r = bicgstab(*(Oracle*)o, rhs, x0, x, tol_abs, tol_rel, max_restarts, max_iter, hist);
  iters_restarts[0] = r.iters;
  iters_restarts[1] = r.restarts;
  errs[0] = r.err_init;
  errs[1] = r.err_opt;
  return r.status;
}

// Pipelined BiCGSTAB with block-CG preconditioner (PoissonSolverAMR.cpp:13-358).
int orc_pipelined(void* o, const double* rhs, const double* x0, double* x, double tol_abs,
                  double tol_rel, int max_iter, int* iters_restarts, double* errs) {
  BiRes r = pipelined(*(Oracle*)o, rhs, x0, x, tol_abs, tol_rel, max_iter);
  iters_restarts[0] = r.iters;
  iters_restarts[1] = r.restarts;
  errs[0] = r.err_init;
  errs[1] = r.err_opt;
  return r.status;
}

// KernelPressureRHS (vel, udef [n][3]; chi [n]; udef/chi may be null).
void orc_pressure_rhs(void* o, const double* vel, const double* udef, const double* chi,
                      double dt, double* out) {
  pressure_rhs(((Oracle*)o)->g, vel, udef, chi, dt, out);
}
// KernelGradP + velocity update, vel [n][3] updated in place.
void orc_grad_p_update(void* o, const double* p, double dt, double* vel) {
  grad_p_update(((Oracle*)o)->g, p, dt, vel);
}

// Implicit diffusion (test infrastructure): operator of velocity component
// `comp` with sigma = 1 / (nu dt) (sigma = 0: h lap), on a scalar field.
void orc_diffusion_lhs(void* o, int comp, double sigma, const double* x, double* y) {
  diffusion_lhs(((Oracle*)o)->g, comp, sigma, x, 1, 0, y);
}
// The same on component `comp` of a [n][3] field (the rhs Laplacian).
void orc_diffusion_lap_vec(void* o, int comp, const double* vel, double* y) {
  diffusion_lhs(((Oracle*)o)->g, comp, 0.0, vel, 3, comp, y);
}
// Exact per-level block inverse and the reference's block CG.
void orc_diffusion_precond(void* o, double sigma, const double* p, double* z) {
  DiffusionPinv D;
  diffusion_precond(*(Oracle*)o, sigma, D, p, z);
}
void orc_diffusion_block_cg(void* o, double sigma, const double* p, double* z) {
  block_cg(*(Oracle*)o, p, z, sigma);
}
// Classical BiCGSTAB (BiCGSTAB.cu:416-628) on the diffusion system with the
// exact per-level block inverse.
int orc_diffusion_bicgstab(void* o, int comp, double sigma, const double* rhs, const double* x0,
                           double* x, double tol_abs, double tol_rel, int max_restarts,
                           int max_iter, int* iters_restarts, double* errs, double* hist) {
  const Oracle& O = *(Oracle*)o;
  const int n = O.M.n;
  DiffusionPinv D;
  std::vector<double> b(rhs, rhs + n), xv(n, 0.);
  if (x0) std::copy(x0, x0 + n, xv.begin());
  BiRes r = bicgstab_core(
      n, b, std::move(xv), x,
      [&](const double* a, double* y) { diffusion_lhs(O.g, comp, sigma, a, 1, 0, y); },
      [&](const double* a, double* y) { diffusion_precond(O, sigma, D, a, y); }, tol_abs,
      tol_rel, max_restarts, max_iter, hist);
  iters_restarts[0] = r.iters;
  iters_restarts[1] = r.restarts;
  errs[0] = r.err_init;
  errs[1] = r.err_opt;
  return r.status;
}
// DiffusionSolver::solve (pipelined, block CG, unlimited restarts).
int orc_diffusion_pipelined(void* o, int comp, double sigma, const double* rhs, const double* x0,
                            double* x, double tol_abs, double tol_rel, int max_iter,
                            int* iters_restarts, double* errs) {
  const Oracle& O = *(Oracle*)o;
  const size_t n = O.M.n;
  std::vector<double> b(rhs, rhs + n), xv(n, 0.);
  if (x0) std::copy(x0, x0 + n, xv.begin());
  BiRes r = pipelined_core(
      n, b, std::move(xv), x,
      [&](const double* a, double* y) { diffusion_lhs(O.g, comp, sigma, a, 1, 0, y); },
      [&](const double* a, double* y) { block_cg(O, a, y, sigma); }, tol_abs, tol_rel, max_iter,
      1 << 30);
  iters_restarts[0] = r.iters;
  iters_restarts[1] = r.restarts;
  errs[0] = r.err_init;
  errs[1] = r.err_opt;
  return r.status;
}
// AdvectionDiffusionImplicit::euler after the advection kernel; iters[3],
// errs[6] = (err_init, err_opt) per component.
void orc_diffusion_step(void* o, double nu, double dt, double tol_abs, double tol_rel,
                        int max_iter, const double* vel_old, double* vel, int* iters,
                        double* errs) {
  diffusion_step(*(Oracle*)o, nu, dt, tol_abs, tol_rel, max_iter, vel_old, vel, iters, errs);
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
