// Internal declarations shared by the host side (topology, solver driver,
// C ABI) and the HIP kernels of libcup3d_poisson.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace cup3d {

constexpr int BS = 8;          // cells per block edge (Definitions.h:26-30)
constexpr int BLEN = 512;      // cells per block
constexpr int WG_THREADS = 256;

// Face classes of a block face, resolved once per topology.
enum FaceKind : int8_t { FACE_NONE = 0, FACE_SAME = 1, FACE_COARSE = 2, FACE_FINE = 3 };

// Status codes written by the device state machine (and returned by solve).
enum SolveStatus : int {
  ST_RUNNING = -1,
  ST_CONVERGED = 0,
  ST_MAXITER = 1,
  ST_BREAKDOWN = 2,
  ST_NONFINITE = 3,
};

// Device-resident solver scalars. The first group mirrors the reference's
// BiCGSTABScalars (source/poisson/BiCGSTAB.cuh:12-21) and its init
// {1, 1, 1, 1e-21, 1, 1, 0, 0} (BiCGSTAB.cu:431); the rest is the host-side
// control of BiCGSTABSolver::main (:424-428, :474-607) moved on device so no
// host round trip is needed per iteration.
struct alignas(64) SolverState {
  double alpha, beta, omega, eps, rho_prev, rho_curr;
  double rr;        // ||r||^2 of the current residual
  double rhrh;      // ||rhat||^2
  double dot_a;     // (rhat, nu)
  double dot_ts;    // (t, s)
  double dot_tt;    // (t, t)
  double err, err_init, err_opt;
  double tol_abs, tol_rel;
  double nu2;       // ||A x0||^2 (reported)
  double bnorm2;    // ||b||^2 (reported)
  double mean_sum;  // sum_b h_b^3 sum x of the last operator application
  double true_rr;   // ||b - A x||^2 of the last residual evaluation
  int iter;         // iterations started
  int iters_done;   // iterations completed
  int restarts;
  int max_restarts;
  int max_iter;
  int status;       // SolveStatus
  int restart_pending;  // K1 copies rhat = r
  int p_zero;           // K1 uses p = r (p = nu = 0)
  int xcur, xopt;       // rotating iterate buffers (0..2)
  int pad[2];
};

// Workspace passed to the kernels (all device pointers).
template <typename T>
struct KrylovVectors {
  T* x[3];   // rotating iterate buffers
  T* r;      // residual (also s)
  T* rhat;
  T* p;
  T* nu;
  T* t;
  T* z1;     // M^-1 p
  T* z2;     // M^-1 s
  T* b;
};

// Grid data on the device.
template <typename T>
struct DevGrid {
  int nb = 0;
  long long n = 0;
  const int* nbr = nullptr;         // [nb][6] same-level neighbour id, or -1 (ghost = self)
  const T* h = nullptr;             // [nb] cell size
  const T* invh = nullptr;          // [nb]
  const double* h3 = nullptr;       // [nb] h^3 (always FP64: mean constraint)
  // coarse-fine faces, applied matrix-free by the SpMV: kind of face f of
  // block b (FaceKind), its partners (coarse: fnb[4(6b+f)]; finer: the four
  // children fnb[4(6b+f) + q], q = (t1 half) + 2 (t2 half)), and the parity
  // of the block index (bit a = index[a] odd) that places a fine block in
  // its coarse neighbour
  const int8_t* fkind = nullptr;    // [nb*6]
  const int* fnb = nullptr;         // [nb*24]
  const int8_t* bpar = nullptr;     // [nb]
  // the same per block in one 128-byte record: fnb (24), fkind (6), bpar, 0
  const int* bmeta = nullptr;       // [nb][32]
  long long bmean_row = -1;         // cell replaced by sum h^3 x (bMean == 1)
  int bmean = 1;
  int has_iface = 0;                // any coarse-fine interface rows
  // implicit diffusion of one velocity component (FP64 only): the operator
  // gains -h^3 helm_sigma x (helm_sigma = 1 / (nu dt)) and, on the domain-
  // boundary faces whose bit f is set in dmask, the Dirichlet ghost -x
  // (h (-x - x) per face) instead of the zero-flux ghost x
  int helm = 0;
  int dmask = 0;
  double helm_sigma = 0.0;
};

// Host-side topology build output.
struct HostTopology {
  int nb = 0;
  std::vector<int> nbr;             // nb*6
  std::vector<int8_t> kind;         // nb*6
  std::vector<double> h, invh, h3;
  std::vector<int> irow_ptr, irow_cell, irow_ent, ent_col;  // CSR form of the interface rows
  std::vector<double> ent_val;                              // (debug matrix only)
  std::vector<int> fnb;             // nb*24 coarse/fine partners per face
  std::vector<int8_t> bpar;         // nb index parities
  long long bmean_row = -1;
  int origin_block = -1;
  long long n_coarse_faces = 0, n_fine_faces = 0;
};

struct TopologyInput {
  int bpd[3];
  int level_max;
  int bc[3];       // 0 freespace, 1 periodic, 2 wall
  double h0;
  int bmean;
  int nb;
  const int* blocks;  // nb x 4 (level, ix, iy, iz)
  bool rows = false;  // also assemble the explicit interface rows (debug matrix only)
};

// The message cup3d_poisson_last_error(NULL) returns (entry points without
// a handle).
void set_global_error(const std::string& m);

// topology.cpp
bool build_topology(const TopologyInput& in, HostTopology& out, std::string& err);
void build_pinv_dst(std::vector<double>& P);  // 512x512 row-major, = A_loc^-1 with sign of the reference

// Precond GEMM operand packing (kernels_precond.hip)
void pack_pinv_f64(const std::vector<double>& P, std::vector<double>& packed);

}  // namespace cup3d
