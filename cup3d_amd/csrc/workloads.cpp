// Synthetic Poisson workloads: block lists (uniform and 2:1-balanced AMR
// octrees) and right-hand sides in Cubism's 8^3-cell block layout.
//
// These are the inputs a CubismUP_3D time step would hand to
// PoissonSolverBase::solve() (rhs in sim.lhs, guess in sim.pres), built
// without the fluid solver:
//   * rhs = discrete divergence of a penalised velocity field, with the
//     volume weighting of KernelPressureRHS
//     (source/operators/PressureProjection.cpp:114-141):
//        b = h^2/(2 dt) * (uE - uW + vN - vS + wB - wF),  dt = 1,
//     where u = (1 - chi) * sum_b A_b exp(-|x - c_b|^2 / (2 sigma_b^2)) is
//     evaluated analytically at the six neighbouring cell centres (so
//     coarse-fine faces need no flux correction), and chi = 1 inside the
//     obstacles, 0 outside; velocity ghosts outside a wall are zero;
//   * or rhs = h^3 * sum_m a_m cos(2 pi k_m . x + phi_m) (periodic modes);
//   * then the volume-weighted mean is removed, b_i -= h_i^3 sum(b)/sum(h^3),
//     the compatibility projection of PressureProjection.cpp:404-431 applied
//     to the rhs.
// Random numbers: std::mt19937_64, uniforms from the top 53 bits, normals by
// Box-Muller written out here (std::normal_distribution is
// implementation-defined, so it is not used).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr int BS = 8;

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double uni() { return double(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uni(double a, double b) { return a + (b - a) * uni(); }
  double normal() {  // Box-Muller, one value per call (the second is discarded)
    double u1 = uni();
    while (u1 <= 0.0) u1 = uni();
    const double u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct Blob {
  double c[3], d[3], sigma, A[3];
};
struct Ellipsoid {
  double c[3], ax[3], R[3][3];  // R rows: body axes in world frame
};
struct Scene {
  double ext[3];
  int wall[3];  // 1 = wall (zero ghost velocity), 0 = periodic (wrap)
  std::vector<Blob> blobs;
  double speed = 0.005;
  int step = 0;
  bool sphere = false;
  double sc[3], sr;
  std::vector<Ellipsoid> ells;
};

inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

struct Leaf {
  int l, i, j, k;
};
inline uint64_t lkey(int l, int i, int j, int k) {
  return (uint64_t(l) << 60) | (uint64_t(i) << 40) | (uint64_t(j) << 20) | uint64_t(k);
}

// Octree order: sort by the Morton code of the block's lowest corner at the
// finest level, ties (impossible for leaves) broken by level.
void sort_leaves(std::vector<Leaf>& L, int lmax) {
  std::sort(L.begin(), L.end(), [lmax](const Leaf& a, const Leaf& b) {
    const int sa = lmax - a.l, sb = lmax - b.l;
    const uint64_t ma = morton3(a.i << sa, a.j << sa, a.k << sa);
    const uint64_t mb = morton3(b.i << sb, b.j << sb, b.k << sb);
    return ma != mb ? ma < mb : a.l < b.l;
  });
}

// Distance from point to ellipsoid surface, first-order estimate |q-1|/|grad q|
// with q = |R(x-c)/axes|.
double ell_dist(const Ellipsoid& e, const double x[3]) {
  double y[3];
  for (int a = 0; a < 3; ++a) {
    y[a] = 0;
    for (int b = 0; b < 3; ++b) y[a] += e.R[a][b] * (x[b] - e.c[b]);
  }
  double q2 = 0;
  for (int a = 0; a < 3; ++a) q2 += (y[a] / e.ax[a]) * (y[a] / e.ax[a]);
  const double q = std::sqrt(q2);
  if (q < 1e-12) return std::min(e.ax[0], std::min(e.ax[1], e.ax[2]));
  double g2 = 0;
  for (int a = 0; a < 3; ++a) {
    const double ga = y[a] / (e.ax[a] * e.ax[a]) / q;
    g2 += ga * ga;
  }
  return std::fabs(q - 1.0) / std::sqrt(g2);
}
bool in_ellipsoid(const Ellipsoid& e, const double x[3]) {
  double q2 = 0;
  for (int a = 0; a < 3; ++a) {
    double y = 0;
    for (int b = 0; b < 3; ++b) y += e.R[a][b] * (x[b] - e.c[b]);
    q2 += (y / e.ax[a]) * (y / e.ax[a]);
  }
  return q2 <= 1.0;
}

double chi_at(const Scene& S, const double x[3], const std::vector<const Ellipsoid*>& ells) {
  if (S.sphere) {
    const double dx = x[0] - S.sc[0], dy = x[1] - S.sc[1], dz = x[2] - S.sc[2];
    if (dx * dx + dy * dy + dz * dz <= S.sr * S.sr) return 1.0;
  }
  for (const Ellipsoid* e : ells)
    if (in_ellipsoid(*e, x)) return 1.0;
  return 0.0;
}

// Ellipsoids whose bounding sphere meets the box [lo, hi].
std::vector<const Ellipsoid*> near_ells(const Scene& S, const double lo[3], const double hi[3]) {
  std::vector<const Ellipsoid*> v;
  for (const auto& e : S.ells) {
    const double r = std::max(e.ax[0], std::max(e.ax[1], e.ax[2]));
    double d2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double q = std::max(lo[a] - e.c[a], std::max(0.0, e.c[a] - hi[a]));
      d2 += q * q;
    }
    if (d2 <= r * r) v.push_back(&e);
  }
  return v;
}

// Velocity of the penalised blob field at point x.
void vel_at(const Scene& S, const double xin[3], const std::vector<const Blob*>& near,
            const std::vector<const Ellipsoid*>& ells, double u[3]) {
  double x[3];
  u[0] = u[1] = u[2] = 0;
  for (int a = 0; a < 3; ++a) {
    x[a] = xin[a];
    if (x[a] < 0 || x[a] > S.ext[a]) {
      if (S.wall[a]) return;  // zero velocity beyond a wall
      x[a] = std::fmod(std::fmod(x[a], S.ext[a]) + S.ext[a], S.ext[a]);
    }
  }
  if (chi_at(S, x, ells) > 0) return;
  for (const Blob* b : near) {
    double r2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double c = b->c[a] + S.step * S.speed * b->d[a];
      r2 += (x[a] - c) * (x[a] - c);
    }
    const double g = std::exp(-r2 / (2 * b->sigma * b->sigma));
    for (int a = 0; a < 3; ++a) u[a] += b->A[a] * g;
  }
}

void remove_mean(int nb, const int* blocks, double h0, double* b) {
  double s = 0, v = 0;
  for (int i = 0; i < nb; ++i) {
    const double h = h0 / double(1 << blocks[4 * i]), h3 = h * h * h;
    double bs = 0;
    for (int c = 0; c < 512; ++c) bs += b[(size_t)i * 512 + c];
    s += bs;
    v += 512 * h3;
  }
  const double m = s / v;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    const double h = h0 / double(1 << blocks[4 * i]), h3 = h * h * h;
    for (int c = 0; c < 512; ++c) b[(size_t)i * 512 + c] -= h3 * m;
  }
}

// Blobs whose 8-sigma ball meets the box [lo, hi] at the scene's step.
std::vector<const Blob*> near_blobs(const Scene& S, const double lo[3], const double hi[3]) {
  std::vector<const Blob*> near;
  for (const auto& b : S.blobs) {
    double d2 = 0;
    for (int a = 0; a < 3; ++a) {
      const double c = b.c[a] + S.step * S.speed * b.d[a];
      const double q = std::max(lo[a] - c, std::max(0.0, c - hi[a]));
      d2 += q * q;
    }
    const double rc = 8 * b.sigma;
    if (d2 <= rc * rc) near.push_back(&b);
  }
  return near;
}

// The penalised blob velocity at the cell centres, [n][3] (Cubism's
// VectorBlock layout: [block][iz][iy][ix][component]).
void blob_vel(const Scene& S, int nb, const int* blocks, double h0, double* out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nb; ++i) {
    const double h = h0 / double(1 << blocks[4 * i]);
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = blocks[4 * i + 1 + a] * BS * h;
      hi[a] = lo[a] + BS * h;
    }
    const auto near = near_blobs(S, lo, hi);
    const auto ells = near_ells(S, lo, hi);
    double* o = out + (size_t)i * 512 * 3;
    for (int c = 0; c < 512; ++c) {
      const double x[3] = {lo[0] + ((c & 7) + 0.5) * h, lo[1] + (((c >> 3) & 7) + 0.5) * h,
                           lo[2] + ((c >> 6) + 0.5) * h};
      vel_at(S, x, near, ells, o + 3 * c);
    }
  }
}

void blob_rhs(const Scene& S, int nb, const int* blocks, double h0, double* out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nb; ++i) {
    const int l = blocks[4 * i];
    const double h = h0 / double(1 << l);
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = blocks[4 * i + 1 + a] * BS * h - h;
      hi[a] = lo[a] + (BS + 2) * h;
    }
    // blobs whose 8-sigma ball meets the block's ghost-extended box
    const auto near = near_blobs(S, lo, hi);
    const auto ells = near_ells(S, lo, hi);
    const double fac = 0.5 * h * h;  // h^2/(2 dt), dt = 1
    double* o = out + (size_t)i * 512;
    for (int iz = 0; iz < BS; ++iz)
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          const double x[3] = {lo[0] + (ix + 1.5) * h, lo[1] + (iy + 1.5) * h,
                               lo[2] + (iz + 1.5) * h};
          double div = 0;
          for (int a = 0; a < 3; ++a) {
            double xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
            xp[a] += h;
            xm[a] -= h;
            double up[3], um[3];
            vel_at(S, xp, near, ells, up);
            vel_at(S, xm, near, ells, um);
            div += up[a] - um[a];
          }
          o[iz * 64 + iy * 8 + ix] = fac * div;
        }
  }
  remove_mean(nb, blocks, h0, out);
}

Scene make_blob_scene(const double ext[3], const int bc[3], int nblobs, double smin, double smax,
                      uint64_t seed) {
  Scene S;
  for (int a = 0; a < 3; ++a) {
    S.ext[a] = ext[a];
    S.wall[a] = bc[a] != 1;
  }
  Rng R(seed);
  S.blobs.resize(nblobs);
  for (auto& b : S.blobs) {  // draw order: centre, sigma, amplitude, direction
    for (int a = 0; a < 3; ++a) b.c[a] = R.uni(0, ext[a]);
    b.sigma = (smin == smax) ? smin : R.uni(smin, smax);
    for (int a = 0; a < 3; ++a) b.A[a] = R.normal();
    double n = 0;
    for (int a = 0; a < 3; ++a) {
      b.d[a] = R.normal();
      n += b.d[a] * b.d[a];
    }
    n = std::sqrt(n);
    for (int a = 0; a < 3; ++a) b.d[a] /= n;
  }
  return S;
}

std::vector<Ellipsoid> make_school(int n, const double ext[3], const double ax[3],
                                   uint64_t seed) {
  Rng R(seed);
  std::vector<Ellipsoid> E;
  const double rmax = std::max(ax[0], std::max(ax[1], ax[2]));
  int guard = 0;
  while ((int)E.size() < n && guard++ < 1000000) {
    Ellipsoid e;
    for (int a = 0; a < 3; ++a) e.c[a] = R.uni(0.15, ext[a] - 0.15);
    double q[4], qn = 0;
    for (int a = 0; a < 4; ++a) {
      q[a] = R.normal();
      qn += q[a] * q[a];
    }
    qn = std::sqrt(qn);
    for (int a = 0; a < 4; ++a) q[a] /= qn;
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double M[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                            {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                            {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
    for (int a = 0; a < 3; ++a) {
      e.ax[a] = ax[a];
      for (int b = 0; b < 3; ++b) e.R[a][b] = M[b][a];  // world->body = M^T
    }
    bool ok = true;  // non-overlap: bounding spheres kept apart
    for (const auto& o : E) {
      double d2 = 0;
      for (int a = 0; a < 3; ++a) d2 += (o.c[a] - e.c[a]) * (o.c[a] - e.c[a]);
      if (d2 < 4 * rmax * rmax * 1.1) {
        ok = false;
        break;
      }
    }
    if (ok) E.push_back(e);
  }
  return E;
}

// Refine-until-criterion + 2:1 balance over faces, edges and corners.
// target(level, centre) returns the level the block at that centre wants.
template <class F>
std::vector<Leaf> build_amr(const int bpd[3], double h0, int lmax, F target) {
  std::unordered_set<uint64_t> leaves;
  std::vector<Leaf> L;
  for (int k = 0; k < bpd[2]; ++k)
    for (int j = 0; j < bpd[1]; ++j)
      for (int i = 0; i < bpd[0]; ++i) leaves.insert(lkey(0, i, j, k));
  auto centre = [&](int l, int i, int j, int k, double c[3]) {
    const double bh = BS * h0 / double(1 << l);
    c[0] = (i + 0.5) * bh;
    c[1] = (j + 0.5) * bh;
    c[2] = (k + 0.5) * bh;
  };
  auto refine = [&](uint64_t key) {
    const int l = int(key >> 60), i = int((key >> 40) & 0xfffff), j = int((key >> 20) & 0xfffff),
              k = int(key & 0xfffff);
    leaves.erase(key);
    for (int c = 0; c < 8; ++c)
      leaves.insert(lkey(l + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + ((c >> 2) & 1)));
  };
  for (int l = 0; l < lmax; ++l) {
    std::vector<uint64_t> todo;
    for (uint64_t key : leaves) {
      const int ll = int(key >> 60);
      if (ll != l) continue;
      double c[3];
      centre(l, int((key >> 40) & 0xfffff), int((key >> 20) & 0xfffff), int(key & 0xfffff), c);
      if (target(c) > l) todo.push_back(key);
    }
    for (uint64_t key : todo) refine(key);
  }
  // 2:1 balance: a leaf at level l whose 26-neighbourhood holds a leaf at
  // level > l+1 is refined; repeat until stable.
  // maxdesc[(l, I)] = deepest leaf level strictly inside block (l, I)
  std::unordered_map<uint64_t, int> maxdesc;
  auto covered_by_finer = [&](int l, int i, int j, int k, int lvl_needed) {
    auto it = maxdesc.find(lkey(l, i, j, k));
    return it != maxdesc.end() && it->second >= lvl_needed;
  };
  bool changed = true;
  while (changed) {
    changed = false;
    maxdesc.clear();
    for (uint64_t key : leaves) {
      const int lf = int(key >> 60), i = int((key >> 40) & 0xfffff),
                j = int((key >> 20) & 0xfffff), k = int(key & 0xfffff);
      for (int l = 0; l < lf; ++l) {
        const int s = lf - l;
        int& m = maxdesc[lkey(l, i >> s, j >> s, k >> s)];
        m = std::max(m, lf);
      }
    }
    std::vector<uint64_t> todo;
    for (uint64_t key : leaves) {
      const int l = int(key >> 60), i = int((key >> 40) & 0xfffff),
                j = int((key >> 20) & 0xfffff), k = int(key & 0xfffff);
      if (l + 2 > lmax) continue;
      bool need = false;
      for (int dz = -1; dz <= 1 && !need; ++dz)
        for (int dy = -1; dy <= 1 && !need; ++dy)
          for (int dx = -1; dx <= 1 && !need; ++dx) {
            if (!dx && !dy && !dz) continue;
            const int ni = i + dx, nj = j + dy, nk = k + dz;
            if (ni < 0 || nj < 0 || nk < 0 || ni >= (bpd[0] << l) || nj >= (bpd[1] << l) ||
                nk >= (bpd[2] << l))
              continue;
            if (covered_by_finer(l, ni, nj, nk, l + 2)) need = true;
          }
      if (need) todo.push_back(key);
    }
    for (uint64_t key : todo) {
      refine(key);
      changed = true;
    }
  }
  for (uint64_t key : leaves)
    L.push_back({int(key >> 60), int((key >> 40) & 0xfffff), int((key >> 20) & 0xfffff),
                 int(key & 0xfffff)});
  sort_leaves(L, lmax);
  return L;
}

int emit(const std::vector<Leaf>& L, int* out, int cap) {
  const int n = (int)L.size();
  if (out && cap >= n)
    for (int b = 0; b < n; ++b) {
      out[4 * b] = L[b].l;
      out[4 * b + 1] = L[b].i;
      out[4 * b + 2] = L[b].j;
      out[4 * b + 3] = L[b].k;
    }
  return n;
}

}  // namespace

extern "C" {

// Uniform grid of bx*by*bz level-0 blocks in Morton order. Returns the count;
// fills out (count x 4 ints: level, ix, iy, iz) when cap is large enough.
int cw_uniform_blocks(int bx, int by, int bz, int* out, int cap) {
  std::vector<Leaf> L;
  for (int k = 0; k < bz; ++k)
    for (int j = 0; j < by; ++j)
      for (int i = 0; i < bx; ++i) L.push_back({0, i, j, k});
  sort_leaves(L, 0);
  return emit(L, out, cap);
}

// AMR grid refined around a sphere: a block goes to level m when the distance
// from its centre to the sphere surface is below thr[m-1] (thr decreasing).
int cw_amr_sphere_blocks(const int* bpd, double h0, int lmax, const double* thr,
                         const double* centre, double radius, int* out, int cap) {
  auto target = [&](const double c[3]) {
    const double d = std::fabs(std::sqrt((c[0] - centre[0]) * (c[0] - centre[0]) +
                                         (c[1] - centre[1]) * (c[1] - centre[1]) +
                                         (c[2] - centre[2]) * (c[2] - centre[2])) -
                               radius);
    int t = 0;
    for (int m = 1; m <= lmax; ++m)
      if (d < thr[m - 1]) t = m;
    return t;
  };
  return emit(build_amr(bpd, h0, lmax, target), out, cap);
}

// AMR grid refined around a school of n ellipsoids (seeded positions and
// orientations, non-overlapping). Same threshold rule, minimum distance over
// the school. The ellipsoid set is a function of (n, axes, ext, seed) only.
int cw_amr_school_blocks(const int* bpd, double h0, int lmax, const double* thr, int n,
                         const double* axes, uint64_t seed, int* out, int cap) {
  const double ext[3] = {bpd[0] * BS * h0, bpd[1] * BS * h0, bpd[2] * BS * h0};
  const auto E = make_school(n, ext, axes, seed);
  auto target = [&](const double c[3]) {
    double d = 1e300;
    for (const auto& e : E) d = std::min(d, ell_dist(e, c));
    int t = 0;
    for (int m = 1; m <= lmax; ++m)
      if (d < thr[m - 1]) t = m;
    return t;
  };
  return emit(build_amr(bpd, h0, lmax, target), out, cap);
}

// rhs = h^3 sum_m a_m cos(2 pi k_m . x + phi_m), k_m in {kmin..kmax}^3,
// a_m ~ N(0,1), phi_m ~ U[0, 2 pi); volume-weighted mean removed.
void cw_rhs_fourier(int nb, const int* blocks, double h0, int nmodes, int kmin, int kmax,
                    uint64_t seed, double* out) {
  Rng R(seed);
  std::vector<double> a(nmodes), ph(nmodes);
  std::vector<int> k(3 * nmodes);
  for (int m = 0; m < nmodes; ++m) {  // draw order: k (x,y,z), amplitude, phase
    for (int d = 0; d < 3; ++d) k[3 * m + d] = kmin + int(R.uni() * (kmax - kmin + 1));
    a[m] = R.normal();
    ph[m] = R.uni(0, 2 * M_PI);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    const int l = blocks[4 * i];
    const double h = h0 / double(1 << l), h3 = h * h * h;
    for (int c = 0; c < 512; ++c) {
      const int ix = c & 7, iy = (c >> 3) & 7, iz = c >> 6;
      const double x = (blocks[4 * i + 1] * BS + ix + 0.5) * h;
      const double y = (blocks[4 * i + 2] * BS + iy + 0.5) * h;
      const double z = (blocks[4 * i + 3] * BS + iz + 0.5) * h;
      double f = 0;
      for (int m = 0; m < nmodes; ++m)
        f += a[m] * std::cos(2 * M_PI * (k[3 * m] * x + k[3 * m + 1] * y + k[3 * m + 2] * z) +
                             ph[m]);
      out[(size_t)i * 512 + c] = h3 * f;
    }
  }
  remove_mean(nb, blocks, h0, out);
}

// rhs = divergence of the penalised Gaussian-blob velocity at sequence step
// `step` (centres move `speed` per step along seeded unit directions).
// Obstacles: sphere (if sr > 0) and/or a school of n_ell ellipsoids (seeded
// with ell_seed, as in cw_amr_school_blocks).
void cw_rhs_blobs(int nb, const int* blocks, double h0, const int* bpd, const int* bc,
                  int nblobs, double smin, double smax, uint64_t seed, int step, double speed,
                  const double* sphere_c, double sr, int n_ell, const double* ell_axes,
                  uint64_t ell_seed, double* out) {
  const double ext[3] = {bpd[0] * BS * h0, bpd[1] * BS * h0, bpd[2] * BS * h0};
  Scene S = make_blob_scene(ext, bc, nblobs, smin, smax, seed);
  S.step = step;
  S.speed = speed;
  if (sr > 0) {
    S.sphere = true;
    for (int a = 0; a < 3; ++a) S.sc[a] = sphere_c[a];
    S.sr = sr;
  }
  if (n_ell > 0) S.ells = make_school(n_ell, ext, ell_axes, ell_seed);
  blob_rhs(S, nb, blocks, h0, out);
}

// The penalised blob velocity itself at the cell centres ([nb*512][3],
// VectorBlock layout), same parameters as cw_rhs_blobs: the velocity field a
// time step's implicit diffusion acts on.
void cw_vel_blobs(int nb, const int* blocks, double h0, const int* bpd, const int* bc,
                  int nblobs, double smin, double smax, uint64_t seed, int step, double speed,
                  const double* sphere_c, double sr, int n_ell, const double* ell_axes,
                  uint64_t ell_seed, double* out) {
  const double ext[3] = {bpd[0] * BS * h0, bpd[1] * BS * h0, bpd[2] * BS * h0};
  Scene S = make_blob_scene(ext, bc, nblobs, smin, smax, seed);
  S.step = step;
  S.speed = speed;
  if (sr > 0) {
    S.sphere = true;
    for (int a = 0; a < 3; ++a) S.sc[a] = sphere_c[a];
    S.sr = sr;
  }
  if (n_ell > 0) S.ells = make_school(n_ell, ext, ell_axes, ell_seed);
  blob_vel(S, nb, blocks, h0, out);
}

// The blob parameters cw_rhs_blobs draws for (ext, bc, nblobs, sigma range,
// seed), 10 doubles per blob: centre[3] (at step 0), direction[3], sigma,
// amplitude[3]. For checking the rhs against the analytic divergence.
void cw_blob_params(const double* ext, const int* bc, int nblobs, double smin, double smax,
                    uint64_t seed, double* out) {
  const Scene S = make_blob_scene(ext, bc, nblobs, smin, smax, seed);
  for (int i = 0; i < nblobs; ++i) {
    const Blob& b = S.blobs[i];
    double* o = out + 10 * (size_t)i;
    for (int a = 0; a < 3; ++a) {
      o[a] = b.c[a];
      o[3 + a] = b.d[a];
      o[7 + a] = b.A[a];
    }
    o[6] = b.sigma;
  }
}

// chi (1 inside obstacles, 0 outside) at cell centres, for diagnostics.
void cw_chi(int nb, const int* blocks, double h0, const int* bpd, const double* sphere_c,
            double sr, int n_ell, const double* ell_axes, uint64_t ell_seed, double* out) {
  const double ext[3] = {bpd[0] * BS * h0, bpd[1] * BS * h0, bpd[2] * BS * h0};
  Scene S;
  if (sr > 0) {
    S.sphere = true;
    for (int a = 0; a < 3; ++a) S.sc[a] = sphere_c[a];
    S.sr = sr;
  }
  if (n_ell > 0) S.ells = make_school(n_ell, ext, ell_axes, ell_seed);
  std::vector<const Ellipsoid*> all;
  for (const auto& e : S.ells) all.push_back(&e);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nb; ++i) {
    const double h = h0 / double(1 << blocks[4 * i]);
    for (int c = 0; c < 512; ++c) {
      const double x[3] = {(blocks[4 * i + 1] * BS + (c & 7) + 0.5) * h,
                           (blocks[4 * i + 2] * BS + ((c >> 3) & 7) + 0.5) * h,
                           (blocks[4 * i + 3] * BS + (c >> 6) + 0.5) * h};
      out[(size_t)i * 512 + c] = chi_at(S, x, all);
    }
  }
}

}  // extern "C"
