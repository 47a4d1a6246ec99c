// Host-side topology build: block neighbour classes and the coarse-fine
// interface rows of the AMR operator, plus the block-Jacobi inverse.
//
// The operator is the one PoissonSolverExp assembles as an explicit COO
// matrix (source/poisson/PoissonSolverExp.cpp:221-373). Here it is split in
// two parts that the SpMV kernel applies matrix-free:
//   * the 7-point stencil h * (sum of 6 neighbours - 6 centre), where a face
//     to a same-level block reads that block's face cells, and every other
//     face (domain wall, coarser or finer block) uses ghost = centre, i.e.
//     zero flux (.cpp:319-363);
//   * for cells on coarse-fine faces, an explicit "interface row" holding the
//     flux terms of makeFlux (.cpp:177-217), merged per cell.
// The interface weights come from the same formulas as the reference
// (interpolate, D1, D2, Dmixed: .cpp:112-155, .h:384-478) but are built in
// tensor-product form: the Taylor ghost of a fine cell over the coarse face
// is
//   T = c + s1*D1_t1 + s2*D1_t2 + D2_t1 + D2_t2 + s1*s2*(M_t1 (x) M_t2)
// where t1, t2 are the tangential directions (X faces: y, z; Y faces: z, x;
// Z faces: x, y — .h:158,230,302), each picking a central, backward (coarse
// index 3 or 7) or forward (0 or 4) one-dimensional stencil, and the mixed
// term factors into M_t1 (x) M_t2 with M = {-1/8, 0, 1/8} central, {-1/4, 1/4}
// backward at offsets (-1, 0) and {-1/4, 1/4} forward at (0, 1); this
// factorisation reproduces all nine cases of Dmixed exactly.

#include <algorithm>
#include <cstring>
#include <cmath>
#include <map>
#include <unordered_map>

#include "poisson_internal.h"

namespace cup3d {
namespace {

inline uint64_t key(int l, int i, int j, int k) {
  return (uint64_t(l) << 60) | (uint64_t(i) << 40) | (uint64_t(j) << 20) | uint64_t(k);
}
inline int cell(int ix, int iy, int iz) { return iz * 64 + iy * 8 + ix; }

// One-dimensional stencils over offsets -2..2 (index o+2).
struct Stencil1D {
  double w[5];
};
enum { CEN = 0, BWD = 1, FWD = 2 };
// first derivative, scaled as in D1 (.h:392-404)
const Stencil1D kD1[3] = {{{0, -1. / 8., 0, 1. / 8., 0}},
                          {{1. / 8., -1. / 2., 3. / 8., 0, 0}},
                          {{0, 0, -3. / 8., 1. / 2., -1. / 8.}}};
// second derivative, scaled as in D2 (.h:415-427)
const Stencil1D kD2[3] = {{{0, 1. / 32., -1. / 16., 1. / 32., 0}},
                          {{1. / 32., -1. / 16., 1. / 32., 0, 0}},
                          {{0, 0, 1. / 32., -1. / 16., 1. / 32.}}};
// factor of the mixed derivative (.h:432-477)
const Stencil1D kM[3] = {{{0, -1. / 8., 0, 1. / 8., 0}},
                         {{0, -1. / 4., 1. / 4., 0, 0}},
                         {{0, 0, -1. / 4., 1. / 4., 0}}};

inline int kind_of(int t) {  // coarse tangential index -> stencil class (isBD/isFD)
  if (t == BS - 1 || t == BS / 2 - 1) return BWD;
  if (t == 0 || t == BS / 2) return FWD;
  return CEN;
}

// tangential directions of a face normal to axis d
inline void tangents(int d, int& t1, int& t2) {
  if (d == 0) { t1 = 1; t2 = 2; }
  else if (d == 1) { t1 = 2; t2 = 0; }
  else { t1 = 0; t2 = 1; }
}

// 5x5 Taylor ghost weights around the coarse cell for sign pair (s1, s2).
void taylor_patch(int k1, int k2, double s1, double s2, double W[5][5]) {
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b < 5; ++b) {
      const double da = a == 2, db = b == 2;
      W[a][b] = da * db + s1 * kD1[k1].w[a] * db + s2 * da * kD1[k2].w[b] + kD2[k1].w[a] * db +
                da * kD2[k2].w[b] + s1 * s2 * kM[k1].w[a] * kM[k2].w[b];
    }
}

}  // namespace

bool build_topology(const TopologyInput& in, HostTopology& T, std::string& err) {
  const int nb = in.nb;
  T = HostTopology();
  T.nb = nb;
  std::unordered_map<uint64_t, int> leaf;
  leaf.reserve(size_t(nb) * 2);
  for (int b = 0; b < nb; ++b) {
    const int* B = in.blocks + 4 * b;
    if (B[0] < 0 || B[0] > 15 || B[1] < 0 || B[2] < 0 || B[3] < 0) {
      err = "block " + std::to_string(b) + ": invalid level/index";
      return false;
    }
    for (int d = 0; d < 3; ++d) {
      if (B[1 + d] >= (in.bpd[d] << B[0])) {
        err = "block " + std::to_string(b) + ": index outside the domain";
        return false;
      }
      if ((in.bpd[d] << B[0]) > (1 << 20)) {  // the key holds 20-bit indices
        err = "block " + std::to_string(b) + ": level too fine (more than 2^20 blocks per direction)";
        return false;
      }
    }
    if (!leaf.emplace(key(B[0], B[1], B[2], B[3]), b).second) {
      err = "duplicate block " + std::to_string(b);
      return false;
    }
  }
  auto find = [&](int l, int i, int j, int k) {
    if (l < 0 || l > 15 || i >= (1 << 20) || j >= (1 << 20) || k >= (1 << 20)) return -1;
    auto it = leaf.find(key(l, i, j, k));
    return it == leaf.end() ? -1 : it->second;
  };
  T.nbr.assign(size_t(nb) * 6, -1);
  T.kind.assign(size_t(nb) * 6, FACE_NONE);
  T.h.resize(nb);
  T.invh.resize(nb);
  T.h3.resize(nb);
  // coarse/fine partner per face: coarse block, or fine child per quadrant
  std::vector<int> coarse(size_t(nb) * 6, -1);
  std::vector<int> fine(size_t(nb) * 6 * 4, -1);
  for (int b = 0; b < nb; ++b) {
    const int* B = in.blocks + 4 * b;
    const int l = B[0];
    const double h = in.h0 / double(1 << l);
    T.h[b] = h;
    T.invh[b] = 1.0 / h;
    T.h3[b] = h * h * h;
    if (B[1] == 0 && B[2] == 0 && B[3] == 0) T.origin_block = b;
    for (int f = 0; f < 6; ++f) {
      const int d = f / 2, s = (f & 1) ? 1 : -1;
      const int n = in.bpd[d] << l;
      int I[3] = {B[1], B[2], B[3]};
      const bool boundary = (s < 0) ? I[d] == 0 : I[d] == n - 1;
      if (boundary && in.bc[d] != 1) continue;  // zero-flux wall (.cpp:361)
      I[d] = ((I[d] + s) % n + n) % n;
      const int same = find(l, I[0], I[1], I[2]);
      if (same >= 0) {
        T.kind[b * 6 + f] = FACE_SAME;
        T.nbr[b * 6 + f] = same;
        continue;
      }
      const int c = find(l - 1, I[0] / 2, I[1] / 2, I[2] / 2);
      if (c >= 0) {
        T.kind[b * 6 + f] = FACE_COARSE;
        coarse[b * 6 + f] = c;
        ++T.n_coarse_faces;
        continue;
      }
      int t1, t2;
      tangents(d, t1, t2);
      for (int q = 0; q < 4; ++q) {  // children adjacent to this face
        int C[3];
        for (int a = 0; a < 3; ++a) C[a] = 2 * I[a];
        C[d] += (s < 0) ? 1 : 0;
        C[t1] += q & 1;
        C[t2] += (q >> 1) & 1;
        const int fb = find(l + 1, C[0], C[1], C[2]);
        if (fb < 0) {
          err = "block " + std::to_string(b) + " face " + std::to_string(f) +
                ": neighbour is neither same-level, coarser nor finer (2:1 balance violated)";
          return false;
        }
        fine[(b * 6 + f) * 4 + q] = fb;
      }
      T.kind[b * 6 + f] = FACE_FINE;
      ++T.n_fine_faces;
    }
  }
  if (in.bmean == 1 && T.origin_block >= 0) T.bmean_row = (long long)T.origin_block * BLEN;

  // Interface rows (the debug matrix only; the solver applies the
  // coarse-fine faces matrix-free): one merged row per cell that touches a
  // coarse-fine face.
  T.irow_ptr.assign(nb + 1, 0);
  T.irow_ent.assign(1, 0);
  for (int b = 0; b < nb && in.rows; ++b) {
    const int* B = in.blocks + 4 * b;
    bool any = false;
    for (int f = 0; f < 6; ++f) any |= T.kind[b * 6 + f] >= FACE_COARSE;
    if (any) {
      std::map<int, std::map<long long, double>> rows;  // cell -> (col -> val)
      const long long base = (long long)b * BLEN;
      for (int f = 0; f < 6; ++f) {
        const int k = T.kind[b * 6 + f];
        if (k < FACE_COARSE) continue;
        const int d = f / 2;
        const bool minus = (f & 1) == 0;
        int t1, t2;
        tangents(d, t1, t2);
        const int nrm = minus ? 0 : BS - 1;
        for (int u = 0; u < BS; ++u)
          for (int v = 0; v < BS; ++v) {
            int c[3];
            c[d] = nrm;
            c[t1] = u;
            c[t2] = v;
            auto& row = rows[cell(c[0], c[1], c[2])];
            const long long self = base + cell(c[0], c[1], c[2]);
            if (k == FACE_COARSE) {
              // fine side (.cpp:177-190): h*(ghost - self),
              // ghost = 2/3 self - 1/5 inward + 8/15 T(coarse)
              const int cb = coarse[b * 6 + f];
              const double h = T.h[b];
              int in_c[3] = {c[0], c[1], c[2]};
              in_c[d] += minus ? 1 : -1;
              row[self] += h * (2.0 / 3.0) - h;
              row[base + cell(in_c[0], in_c[1], in_c[2])] += -h / 5.0;
              int cc[3];  // coarse cell (ix_c etc., .h:132-137 and overrides)
              for (int a = 0; a < 3; ++a)
                cc[a] = (B[1 + a] % 2 == 0) ? c[a] / 2 : c[a] / 2 + BS / 2;
              cc[d] = minus ? BS - 1 : 0;
              const double s1 = (c[t1] % 2 == 0) ? -1.0 : 1.0;  // taylorSign
              const double s2 = (c[t2] % 2 == 0) ? -1.0 : 1.0;
              double W[5][5];
              taylor_patch(kind_of(cc[t1]), kind_of(cc[t2]), s1, s2, W);
              const double tf = h * 8.0 / 15.0;
              for (int a = 0; a < 5; ++a)
                for (int bb = 0; bb < 5; ++bb) {
                  if (W[a][bb] == 0.0) continue;
                  int q[3] = {cc[0], cc[1], cc[2]};
                  q[t1] += a - 2;
                  q[t2] += bb - 2;
                  row[(long long)cb * BLEN + cell(q[0], q[1], q[2])] += tf * W[a][bb];
                }
            } else {
              // coarse side (.cpp:191-217): sum over the 4 fine cells of
              // h_f*(p_close - ghost), ghost = 2/3 p_close - 1/5 p_far + 8/15 T_q
              // the child block holding this coarse cell's half in t1, t2
              // (Zchild, .h:198-199,...), then its 4 cells q (neiFine1..4)
              const int fq = (c[t1] >= BS / 2 ? 1 : 0) + (c[t2] >= BS / 2 ? 2 : 0);
              const int fb = fine[(b * 6 + f) * 4 + fq];
              for (int q = 0; q < 4; ++q) {
                const double hf = T.h[fb];
                int fc[3], ff[3];
                fc[t1] = (c[t1] % (BS / 2)) * 2 + (q & 1);  // ix_f (.h:102-104)
                fc[t2] = (c[t2] % (BS / 2)) * 2 + ((q >> 1) & 1);
                fc[d] = minus ? BS - 1 : 0;
                ff[t1] = fc[t1];
                ff[t2] = fc[t2];
                ff[d] = minus ? BS - 2 : 1;
                const long long fbase = (long long)fb * BLEN;
                row[fbase + cell(fc[0], fc[1], fc[2])] += hf - hf * (2.0 / 3.0);
                row[fbase + cell(ff[0], ff[1], ff[2])] += hf / 5.0;
                const double s1 = (q & 1) ? 1.0 : -1.0, s2 = (q & 2) ? 1.0 : -1.0;
                double W[5][5];
                taylor_patch(kind_of(c[t1]), kind_of(c[t2]), s1, s2, W);
                const double tf = -hf * 8.0 / 15.0;
                for (int a = 0; a < 5; ++a)
                  for (int bb = 0; bb < 5; ++bb) {
                    if (W[a][bb] == 0.0) continue;
                    int p[3] = {c[0], c[1], c[2]};
                    p[t1] += a - 2;
                    p[t2] += bb - 2;
                    row[base + cell(p[0], p[1], p[2])] += tf * W[a][bb];
                  }
              }
            }
          }
      }
      for (auto& kv : rows) {
        T.irow_cell.push_back(kv.first);
        for (auto& cv : kv.second) {
          if (cv.second == 0.0) continue;
          if (cv.first > 0x7fffffffLL) {
            err = "grid too large for 32-bit interface column indices";
            return false;
          }
          T.ent_col.push_back((int)cv.first);
          T.ent_val.push_back(cv.second);
        }
        T.irow_ent.push_back((int)T.ent_col.size());
      }
    }
    T.irow_ptr[b + 1] = (int)T.irow_cell.size();
  }
  // partners and parities for the matrix-free coarse-fine stencils
  T.fnb.assign(size_t(nb) * 24, -1);
  T.bpar.assign(nb, 0);
  for (int b = 0; b < nb; ++b) {
    const int* B = in.blocks + 4 * b;
    T.bpar[b] = int8_t((B[1] & 1) | ((B[2] & 1) << 1) | ((B[3] & 1) << 2));
    for (int f = 0; f < 6; ++f) {
      if (T.kind[b * 6 + f] == FACE_COARSE) T.fnb[(b * 6 + f) * 4] = coarse[b * 6 + f];
      if (T.kind[b * 6 + f] == FACE_FINE)
        for (int q = 0; q < 4; ++q) T.fnb[(b * 6 + f) * 4 + q] = fine[(b * 6 + f) * 4 + q];
    }
  }
  return true;
}

// A_loc^-1 (with the reference's sign: P_inv = -K^-1, K = 6 I - face
// adjacency of the 8^3 block with zero Dirichlet ghosts,
// PoissonSolverExp.cpp:17-35,97-107) from the closed-form eigendecomposition
// of K by the 3-D discrete sine transform:
//   K = sum_abc lambda_abc v_abc v_abc^T,  v_abc = s_a (x) s_b (x) s_c,
//   s_a(i) = sqrt(2/9) sin(pi a (i+1) / 9),
//   lambda_abc = 6 - 2 cos(pi a/9) - 2 cos(pi b/9) - 2 cos(pi c/9).
void build_pinv_dst(std::vector<double>& P) {
  double S[BS][BS], lam1[BS];
  for (int a = 0; a < BS; ++a) {
    lam1[a] = 2.0 - 2.0 * std::cos(M_PI * (a + 1) / 9.0);
    for (int i = 0; i < BS; ++i) S[a][i] = std::sqrt(2.0 / 9.0) * std::sin(M_PI * (a + 1) * (i + 1) / 9.0);
  }
  // separable: P[(i,j,k),(i',j',k')] = -sum_abc Sa(i)Sa(i') Sb(j)Sb(j') Sc(k)Sc(k') / lam
  double Q[BS][BS][BS];  // Q[a][i][i'] = S[a][i] S[a][i']
  for (int a = 0; a < BS; ++a)
    for (int i = 0; i < BS; ++i)
      for (int j = 0; j < BS; ++j) Q[a][i][j] = S[a][i] * S[a][j];
  P.assign(size_t(BLEN) * BLEN, 0.0);
  for (int r = 0; r < BLEN; ++r) {
    const int i = r % 8, j = (r / 8) % 8, k = r / 64;
    for (int c = r; c < BLEN; ++c) {
      const int i2 = c % 8, j2 = (c / 8) % 8, k2 = c / 64;
      double s = 0;
      for (int cz = 0; cz < BS; ++cz) {
        const double qz = Q[cz][k][k2];
        for (int by = 0; by < BS; ++by) {
          const double qyz = qz * Q[by][j][j2];
          for (int ax = 0; ax < BS; ++ax) s += qyz * Q[ax][i][i2] / (lam1[ax] + lam1[by] + lam1[cz]);
        }
      }
      P[size_t(r) * BLEN + c] = -s;
      P[size_t(c) * BLEN + r] = -s;
    }
  }
}

}  // namespace cup3d

// ---------------------------------------------------------------------------
// Host-only export of the operator the SpMV kernel applies, as CSR (sorted
// columns, duplicates merged). Diagnostics for checking the topology build
// against an assembled matrix without a GPU.
// ---------------------------------------------------------------------------
#include "cup3d_poisson.h"
#include "cup3d_poisson_debug.h"

extern "C" int64_t cup3d_debug_operator_csr(const cup3d_poisson_config* cfg, int64_t nblocks,
                                            const cup3d_block_desc* blocks, int64_t* rowptr,
                                            int32_t* col, double* val, int64_t cap) {
  using namespace cup3d;
  if (!cfg || !blocks || nblocks <= 0) return -1;
  std::vector<int> bl(nblocks * 4);
  for (int64_t b = 0; b < nblocks; ++b) {
    bl[4 * b] = blocks[b].level;
    for (int d = 0; d < 3; ++d) bl[4 * b + 1 + d] = blocks[b].index[d];
  }
  TopologyInput in;
  for (int d = 0; d < 3; ++d) {
    in.bpd[d] = cfg->bpd[d];
    in.bc[d] = cfg->bc[d];
  }
  in.level_max = cfg->level_max;
  in.h0 = cfg->h0;
  in.bmean = cfg->bmean;
  in.nb = (int)nblocks;
  in.blocks = bl.data();
  in.rows = true;
  HostTopology T;
  std::string err;
  if (!build_topology(in, T, err)) return -2;
  int64_t nnz = 0;
  std::map<long long, double> row;
  const int nb = (int)nblocks;
  for (int b = 0; b < nb; ++b) {
    const int* nbr = &T.nbr[b * 6];
    int r = T.irow_ptr[b];
    for (int c = 0; c < BLEN; ++c) {
      const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
      const long long self = (long long)b * BLEN + c;
      const double h = T.h[b];
      row.clear();
      if (cfg->bmean == 3 && b == T.origin_block && c == 0) {
        row[self] = h;
      } else {
        const int cc[3] = {ix, iy, iz};
        for (int f = 0; f < 6; ++f) {
          const int d = f / 2, s = (f & 1) ? 1 : -1;
          int q[3] = {ix, iy, iz};
          q[d] += s;
          if (q[d] >= 0 && q[d] < BS) {
            row[(long long)b * BLEN + cell(q[0], q[1], q[2])] += h;
            row[self] -= h;
          } else if (nbr[f] >= 0) {
            q[d] = (q[d] + BS) % BS;
            row[(long long)nbr[f] * BLEN + cell(q[0], q[1], q[2])] += h;
            row[self] -= h;
          }
          (void)cc;
        }
        if (r < T.irow_ptr[b + 1] && T.irow_cell[r] == c) {
          for (int e = T.irow_ent[r]; e < T.irow_ent[r + 1]; ++e) row[T.ent_col[e]] += T.ent_val[e];
        }
      }
      if (r < T.irow_ptr[b + 1] && T.irow_cell[r] == c) ++r;
      for (auto& kv : row) {
        if (nnz < cap) {
          col[nnz] = (int32_t)kv.first;
          val[nnz] = kv.second;
        }
        ++nnz;
      }
      if (rowptr && cap > 0) rowptr[self + 1] = nnz;
    }
  }
  if (rowptr && cap > 0) rowptr[0] = 0;
  return nnz;
}

extern "C" int cup3d_debug_host_topology(const cup3d_poisson_config* cfg, int64_t nblocks,
                                         const cup3d_block_desc* blocks, int32_t* nbr,
                                         int8_t* kind, int32_t* fnb, int8_t* bpar, double* h,
                                         double* invh, double* h3, int64_t* info, char* err,
                                         int errlen) {
  using namespace cup3d;
  auto set_err = [&](const std::string& m) {
    if (err && errlen > 0) {
      const size_t k = std::min(m.size(), (size_t)errlen - 1);
      std::memcpy(err, m.data(), k);
      err[k] = 0;
    }
    return -1;
  };
  if (!cfg || !blocks || nblocks <= 0 || nblocks > (1LL << 22)) return set_err("bad arguments");
  std::vector<int> bl(nblocks * 4);
  for (int64_t b = 0; b < nblocks; ++b) {
    bl[4 * b] = blocks[b].level;
    for (int d = 0; d < 3; ++d) bl[4 * b + 1 + d] = blocks[b].index[d];
  }
  TopologyInput in;
  for (int d = 0; d < 3; ++d) {
    in.bpd[d] = cfg->bpd[d];
    in.bc[d] = cfg->bc[d];
  }
  in.level_max = cfg->level_max;
  in.h0 = cfg->h0;
  in.bmean = cfg->bmean;
  in.nb = (int)nblocks;
  in.blocks = bl.data();
  HostTopology T;
  std::string e;
  if (!build_topology(in, T, e)) return set_err(e);
  const size_t nb = (size_t)nblocks;
  if (nbr) std::memcpy(nbr, T.nbr.data(), nb * 6 * sizeof(int32_t));
  if (kind) std::memcpy(kind, T.kind.data(), nb * 6);
  if (fnb) std::memcpy(fnb, T.fnb.data(), nb * 24 * sizeof(int32_t));
  if (bpar) std::memcpy(bpar, T.bpar.data(), nb);
  if (h) std::memcpy(h, T.h.data(), nb * sizeof(double));
  if (invh) std::memcpy(invh, T.invh.data(), nb * sizeof(double));
  if (h3) std::memcpy(h3, T.h3.data(), nb * sizeof(double));
  if (info) {
    info[0] = T.origin_block;
    info[1] = T.n_coarse_faces;
    info[2] = T.n_fine_faces;
  }
  return 0;
}
