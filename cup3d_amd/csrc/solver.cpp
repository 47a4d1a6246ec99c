// Solver driver and C ABI of libcup3d_poisson (host side, C++17 + HIP runtime).
//
// Mirrors the life cycle of the reference GPU backend:
//   PoissonSolverExp ctor (P_inv, PoissonSolverExp.cpp:37-111) -> create
//   getMat + updateAll on mesh change (.cpp:437-442, BiCGSTAB.cu:143-247)
//                                                   -> set_topology
//   getVec + BiCGSTABSolver::main + scatter (.cpp:420-473, BiCGSTAB.cu:416-628)
//                                                   -> solve / solve_device
// Device memory is allocated once per topology and reused by every solve.

#include "cup3d_poisson.h"
#include "cup3d_poisson_debug.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "cnn.h"
#include "kernels.h"
#include "mixed.h"
#include "poisson_internal.h"
#include "diffusion.h"
#include "topology_gpu.h"

using namespace cup3d;

namespace {

thread_local std::string g_create_error;

}  // namespace

void cup3d::set_global_error(const std::string& m) { g_create_error = m; }

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    release();
    bytes = b;
    return hipMalloc(&p, b ? b : 1);
  }
  // at least b bytes; keeps the allocation when it is large enough (the
  // contents are not preserved either way)
  hipError_t reserve(size_t b) {
    if (p && bytes >= b) return hipSuccess;
    return alloc(b);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

#define HIPCHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      H->err = std::string(#expr) + ": " + hipGetErrorString(e_);                          \
      return -1;                                                                           \
    }                                                                                      \
  } while (0)

}  // namespace

// Device topology tables of one generation (topology_gpu.hip).
struct DevTopo {
  DevBuf blocks, nbr, fkind, fnb, bpar, h, invh, h3, hf, invhf, meta, keys, vals, info;
};

struct cup3d_poisson {
  cup3d_poisson_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;

  // topology: the tables live on the device (built there by the topology
  // kernels, two generations so that a rejected block list leaves the
  // current one in place); the host keeps the block list, h^3 per block,
  // the origin block and the face counts
  HostTopology topo;
  std::vector<int> blocks;  // nb x 4
  int nb = 0;
  long long n = 0;
  DevTopo tp[2];
  int tcur = 0;
  DevGrid<double> g64;
  DevGrid<float> g32;

  // preconditioner
  std::vector<double> P;
  DevBuf d_P64;
  DevBuf d_dst;    // DST block-inverse table (DST_TABLE doubles), for cfg.precond = DST
  DevBuf d_dst32;  // its FP32 rounding (the mixed forms' inner solve) and implicit diffusion

  // FP64 Krylov workspace
  DevBuf d_vec64;  // 11 vectors of n doubles
  KrylovVectors<double> v64{};
  DevBuf d_state, d_part;
  SolverState* h_state = nullptr;  // pinned
  size_t part_len = 0;

  // mixed precision workspace
  MixedWorkspace mixed;
  // learned warm start
  CnnState cnn;

  // staging of the host entry point: pinned rhs + x and their device copies,
  // allocated by the first host solve on a topology and kept
  double* h_stage = nullptr;  // 2 n doubles, pinned
  DevBuf d_stage;             // 2 n doubles
  long long stage_n = 0;

  // warm start: the guesses handed in by the last cfg.extrapolate solves on
  // this topology (the caller's pressures of the previous steps), most
  // recent first in hist[0 .. hist_count); the images A d_k of the time
  // directions of the last solve (d_k of one solve is d_{k-1} of the next,
  // so only A d_1 is new each time); the Gram matrix
  DevBuf d_hist, d_img, d_gdots;
  double* hist[HIST_MAX] = {};
  double* img[HIST_MAX] = {};
  int hist_count = 0;
  // the FP64 true residual of each kept pressure (that of the solve it came
  // from, when the caller handed that solution back), and of the last solve
  double hist_rho[HIST_MAX] = {};
  double last_resid = -1.0;  // < 0: unknown
};

namespace {

int fail(cup3d_poisson* H, const std::string& m) {
  H->err = m;
  return -1;
}

// Device vectors handed to an entry point: non-null and 16-byte aligned (the
// kernels load and store them as 16-byte vectors). Checked on the host before
// any launch.
int check_vectors(cup3d_poisson* H, std::initializer_list<const void*> ptrs) {
  for (const void* p : ptrs) {
    if (!p) return fail(H, "null vector");
    if ((uintptr_t)p & 15) return fail(H, "device vectors must be 16-byte aligned");
  }
  return 0;
}

SolverState make_state(const cup3d_poisson_config& c) {
  SolverState s;
  std::memset(&s, 0, sizeof(s));
  s.alpha = 1.0;  // BiCGSTAB.cu:431
  s.beta = 1.0;
  s.omega = 1.0;
  s.eps = 1e-21;
  s.rho_prev = 1.0;
  s.rho_curr = 1.0;
  s.err_init = s.err_opt = s.err = 1e50;  // :424-426
  s.tol_abs = c.tol_abs;
  s.tol_rel = c.tol_rel;
  s.max_iter = c.max_iter;
  s.max_restarts = c.max_restarts;
  s.status = ST_RUNNING;
  s.p_zero = 1;
  s.xcur = 0;
  s.xopt = 0;
  return s;
}

int upload(cup3d_poisson* H, DevBuf& d, const void* src, size_t bytes) {
  if (d.alloc(bytes) != hipSuccess) return fail(H, "hipMalloc failed (" + std::to_string(bytes) + " bytes)");
  if (bytes && hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(H, "hipMemcpy H2D failed");
  return 0;
}

// M^-1 of the FP64 form: the P_inv GEMM on the FP64 matrix cores, or (Helmholtz
// grids, or cfg.precond = DST) the exact block inverse by sine transforms.
void launch_precond_fp64(cup3d_poisson* H, const DevGrid<double>& g, int mode,
                         const KrylovVectors<double>& v) {
  SolverState* st = H->d_state.as<SolverState>();
  if (g.helm || H->cfg.precond == CUP3D_PRECOND_DST)
    launch_precond_dst(H->stream, mode, H->nb, H->d_dst.as<double>(), g.h,
                       g.helm ? g.helm_sigma : 0.0, v, st);
  else
    launch_precond64(H->stream, mode, H->nb, H->d_P64.as<double>(), g.invh, v, st);
}

// Run the FP64 classical BiCGSTAB on the device. b64/x64 must already hold
// getVec's b and x0 (in v64.b and v64.x[0]); the state must be initialised.
// With a Helmholtz grid (g.helm, implicit diffusion) the preconditioner is
// the DST block inverse of each block's level instead of the P_inv GEMM.
int run_bicgstab64(cup3d_poisson* H, const DevGrid<double>& g) {
  hipStream_t s = H->stream;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  const KrylovVectors<double>& v = H->v64;
  auto precond = [&](int mode) { launch_precond_fp64(H, g, mode, v); };
  // r = b - A x0; rhat = r (BiCGSTAB.cu:435-467)
  launch_spmv64(s, SPMV_INIT, g, v.x[0], v.nu, v.b, v.r, v.rhat, st, part);
  const int poll = H->cfg.poll_every > 0 ? H->cfg.poll_every : 8;
  // Batches of 1, 2, 4, ... iterations up to `poll` between status polls: a
  // short solve (the implicit-diffusion solves take a few iterations) is not
  // held up by the host enqueueing the rest of a full batch, whose kernels
  // would only find the state converged; a long solve polls a few more times.
  int batch = 1;
  long long launched = 0;
  while (true) {
    for (int k = 0; k < batch; ++k) {
      precond(PREC_P);
      launch_spmv64(s, SPMV_DOT1, g, v.z1, v.nu, v.rhat, nullptr, nullptr, st, part);
      precond(PREC_S);
      launch_spmv64(s, SPMV_DOT2, g, v.z2, v.t, v.r, nullptr, nullptr, st, part);
      launch_update64(s, H->n, v, st, part);
    }
    launched += batch;
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(H->h_state, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (H->h_state->status != ST_RUNNING) break;
    if (launched > (long long)H->h_state->max_iter + 4LL * poll)
      return fail(H, "device state machine did not stop");
    batch = std::min(2 * batch, poll);
  }
  return 0;
}

// Runs fn(b0, b1) over [0, nb) split into contiguous ranges on up to 16 host
// threads (one for small grids): the gather and scatter of the host solve
// are memory-bound copies that one core cannot feed at the host's bandwidth.
template <typename F>
void parallel_blocks(int nb, F&& fn) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, nb / 512));
  if (nt <= 1) {
    fn(0, nb);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  const int chunk = (nb + nt - 1) / nt;
  for (int t = 1; t < nt; ++t) {
    const int b0 = t * chunk, b1 = std::min(nb, b0 + chunk);
    if (b0 < b1) pool.emplace_back([&fn, b0, b1] { fn(b0, b1); });
  }
  fn(0, std::min(nb, chunk));
  for (std::thread& th : pool) th.join();
}

}  // namespace

extern "C" {

int cup3d_poisson_create(const cup3d_poisson_config* cfg, cup3d_poisson_t* out) {
  if (!cfg || !out) {
    g_create_error = "null argument";
    return -1;
  }
  *out = nullptr;
  const cup3d_poisson_config& c = *cfg;
  for (int d = 0; d < 3; ++d) {
    if (c.bpd[d] <= 0 || c.bpd[d] >= 65536) { g_create_error = "bpd must be in 1..65535"; return -1; }
    if (c.bc[d] < 0 || c.bc[d] > 2) { g_create_error = "bc must be 0 (freespace), 1 (periodic) or 2 (wall)"; return -1; }
  }
  if (!(c.h0 > 0)) { g_create_error = "h0 must be positive"; return -1; }
  if (c.bmean < 0 || c.bmean > 3) { g_create_error = "bmean must be 0..3"; return -1; }
  if (c.max_iter <= 0) { g_create_error = "max_iter must be positive"; return -1; }
  if (c.precision < 0 || c.precision > 2) { g_create_error = "unknown precision"; return -1; }
  if (c.guess < 0 || c.guess > 2) { g_create_error = "unknown guess mode"; return -1; }
  if (c.extrapolate < 0 || c.extrapolate > HIST_MAX) {
    g_create_error = "extrapolate must be 0.." + std::to_string(HIST_MAX);
    return -1;
  }
  if (c.precond < 0 || c.precond > 1) { g_create_error = "precond must be 0 (GEMM) or 1 (DST)"; return -1; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_create_error = "no HIP device available";
    return -1;
  }
  if (c.device < 0 || c.device >= ndev) { g_create_error = "invalid device ordinal"; return -1; }
  std::unique_ptr<cup3d_poisson> H(new cup3d_poisson());
  H->cfg = c;
  if (H->cfg.inner_tol <= 0) H->cfg.inner_tol = 1e-7;
  if (H->cfg.max_refine <= 0) H->cfg.max_refine = 10;
  H->mixed.split = c.gemm_passes != 1;
  H->mixed.use_dst = c.precond == CUP3D_PRECOND_DST;
  H->device = c.device;
  if (hipSetDevice(c.device) != hipSuccess || hipStreamCreate(&H->stream) != hipSuccess ||
      hipEventCreate(&H->ev0) != hipSuccess || hipEventCreate(&H->ev1) != hipSuccess ||
      hipHostMalloc((void**)&H->h_state, sizeof(SolverState), hipHostMallocDefault) != hipSuccess) {
    g_create_error = "HIP stream/event/pinned allocation failed";
    return -1;
  }
  // block-Jacobi inverse (PoissonSolverExp ctor) and its MFMA packing
  build_pinv_dst(H->P);
  std::vector<double> pk;
  pack_pinv_f64(H->P, pk);
  double dtab[DST_TABLE];
  float dtab32[DST_TABLE];
  dst_tables(dtab, dtab32);
  if (upload(H.get(), H->d_P64, pk.data(), pk.size() * sizeof(double)) != 0 ||
      upload(H.get(), H->d_dst, dtab, sizeof(dtab)) != 0 ||
      upload(H.get(), H->d_dst32, dtab32, sizeof(dtab32)) != 0 ||
      mixed_init_pinv(H->mixed, H->P, H->err) != 0 || cnn_init(H->cnn, H->err) != 0) {
    g_create_error = H->err;
    return -1;
  }
  H->mixed.dst32 = H->d_dst32.as<float>();
  if (H->d_state.alloc(sizeof(SolverState)) != hipSuccess) {
    g_create_error = "hipMalloc(state) failed";
    return -1;
  }
  *out = H.release();
  return 0;
}

int cup3d_poisson_set_topology(cup3d_poisson_t H, int64_t nblocks, const cup3d_block_desc* blocks) {
  if (!H) return -1;
  if (nblocks <= 0 || !blocks) return fail(H, "empty block list");
  if (nblocks > (1LL << 22)) return fail(H, "too many blocks (limit 4194304)");
  HIPCHECK(hipSetDevice(H->device));
  std::vector<int> bl(nblocks * 4);
  for (int64_t b = 0; b < nblocks; ++b) {
    bl[4 * b] = blocks[b].level;
    for (int d = 0; d < 3; ++d) bl[4 * b + 1 + d] = blocks[b].index[d];
  }
  // build the tables on the device into the other generation: a rejected
  // block list leaves the handle as it was
  const size_t nb = (size_t)nblocks;
  size_t cap = 1024;
  while (cap < 2 * nb) cap <<= 1;
  DevTopo& D = H->tp[1 - H->tcur];
  if (D.blocks.reserve(nb * 4 * sizeof(int)) || D.nbr.reserve(nb * 6 * sizeof(int)) ||
      D.fkind.reserve(nb * 6) || D.fnb.reserve(nb * 24 * sizeof(int)) || D.bpar.reserve(nb) ||
      D.h.reserve(nb * sizeof(double)) || D.invh.reserve(nb * sizeof(double)) ||
      D.h3.reserve(nb * sizeof(double)) || D.hf.reserve(nb * sizeof(float)) ||
      D.invhf.reserve(nb * sizeof(float)) || D.meta.reserve(nb * 32 * sizeof(int)) ||
      D.keys.reserve(cap * sizeof(unsigned long long)) ||
      D.vals.reserve(cap * sizeof(int)) || D.info.reserve(TOPO_INFO_WORDS * sizeof(int)))
    return fail(H, "hipMalloc(topology tables) failed");
  hipStream_t s = H->stream;
  HIPCHECK(hipMemcpyAsync(D.blocks.p, bl.data(), nb * 4 * sizeof(int), hipMemcpyHostToDevice, s));
  TopoParams tpar;
  tpar.nb = (int)nb;
  for (int d = 0; d < 3; ++d) {
    tpar.bpd[d] = H->cfg.bpd[d];
    tpar.bc[d] = H->cfg.bc[d];
  }
  tpar.h0 = H->cfg.h0;
  tpar.blocks = D.blocks.as<int>();
  const TopoTable ttab{D.keys.as<unsigned long long>(), D.vals.as<int>(), (unsigned long long)cap - 1};
  const TopoOut tout{D.nbr.as<int>(),  D.fkind.as<int8_t>(), D.fnb.as<int>(),
                     D.bpar.as<int8_t>(), D.h.as<double>(),  D.invh.as<double>(),
                     D.h3.as<double>(), D.hf.as<float>(),  D.invhf.as<float>(),
                     D.meta.as<int>()};
  launch_topology_build(s, tpar, ttab, tout, D.info.as<int>());
  HIPCHECK(hipGetLastError());
  int info[TOPO_INFO_WORDS];
  HIPCHECK(hipMemcpyAsync(info, D.info.p, sizeof(info), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  // the first offending block in list order, with the message of the
  // sequential checks (level/index, domain, index width, duplicate)
  const int NONE = 0x7fffffff;
  const int first = std::min(std::min(info[TI_INVALID], info[TI_OUTSIDE]),
                             std::min(info[TI_TOO_FINE], info[TI_DUPLICATE]));
  if (first != NONE) {
    const std::string b = "block " + std::to_string(first);
    if (first == info[TI_INVALID]) return fail(H, b + ": invalid level/index");
    if (first == info[TI_OUTSIDE]) return fail(H, b + ": index outside the domain");
    if (first == info[TI_TOO_FINE])
      return fail(H, b + ": level too fine (more than 2^20 blocks per direction)");
    return fail(H, "duplicate " + b);
  }
  if (info[TI_BALANCE] != NONE)
    return fail(H, "block " + std::to_string(info[TI_BALANCE] / 6) + " face " +
                       std::to_string(info[TI_BALANCE] % 6) +
                       ": neighbour is neither same-level, coarser nor finer (2:1 balance violated)");
  if (info[TI_FULL] != NONE) return fail(H, "topology hash table full");
  HostTopology Tn;
  Tn.nb = (int)nb;
  Tn.origin_block = info[TI_ORIGIN];
  Tn.n_coarse_faces = info[TI_N_COARSE];
  Tn.n_fine_faces = info[TI_N_FINE];
  Tn.h3.resize(nb);
  for (size_t b = 0; b < nb; ++b) {
    const double h = H->cfg.h0 / double(1 << bl[4 * b]);
    Tn.h3[b] = h * h * h;
  }
  if (H->cfg.bmean == 1 && Tn.origin_block >= 0) Tn.bmean_row = (long long)Tn.origin_block * BLEN;
  H->tcur = 1 - H->tcur;
  H->blocks.swap(bl);
  H->topo = std::move(Tn);
  // from here on a failure (device allocation) leaves the handle without a
  // topology: solves are refused until set_topology succeeds
  struct Commit {
    cup3d_poisson* h;
    bool ok = false;
    ~Commit() {
      if (!ok) h->nb = 0;
    }
  } commit{H};
  HostTopology& T = H->topo;
  const DevTopo& C = H->tp[H->tcur];
  H->nb = (int)nblocks;
  H->n = (long long)nblocks * BLEN;
  const long long origin_row = T.origin_block >= 0 ? (long long)T.origin_block * BLEN : -1;
  auto fill = [&](auto& g, const void* h, const void* invh) {
    using G = std::remove_reference_t<decltype(g)>;
    using TT = std::remove_const_t<std::remove_pointer_t<decltype(G::h)>>;
    g.nb = H->nb;
    g.n = H->n;
    g.nbr = C.nbr.as<int>();
    g.h = static_cast<const TT*>(h);
    g.invh = static_cast<const TT*>(invh);
    g.h3 = C.h3.as<double>();
    g.fkind = C.fkind.as<int8_t>();
    g.fnb = C.fnb.as<int>();
    g.bpar = C.bpar.as<int8_t>();
    g.bmeta = C.meta.as<int>();
    g.bmean = H->cfg.bmean;
    g.has_iface = T.n_coarse_faces + T.n_fine_faces > 0 ? 1 : 0;
    g.bmean_row = (H->cfg.bmean == 1 || H->cfg.bmean == 3) ? origin_row : -1;
  };
  fill(H->g64, C.h.p, C.invh.p);
  fill(H->g32, C.hf.p, C.invhf.p);
  // FP64 Krylov vectors: x[3], r, rhat, p, nu, t, z1, z2, b
  const size_t n = (size_t)H->n;
  if (H->d_vec64.reserve(11 * n * sizeof(double)) != hipSuccess) return fail(H, "hipMalloc(vectors) failed");
  double* base = H->d_vec64.as<double>();
  KrylovVectors<double>& v = H->v64;
  for (int k = 0; k < 3; ++k) v.x[k] = base + k * n;
  v.r = base + 3 * n;
  v.rhat = base + 4 * n;
  v.p = base + 5 * n;
  v.nu = base + 6 * n;
  v.t = base + 7 * n;
  v.z1 = base + 8 * n;
  v.z2 = base + 9 * n;
  v.b = base + 10 * n;
  // up to GUESS_MAXV (GUESS_MAXV + 1) / 2 partials per workgroup of 2048
  // cells (the warm start's Gram matrix)
  H->part_len = (size_t)((H->nb + 3) / 4 + 8) * (GUESS_MAXV * (GUESS_MAXV + 1) / 2);
  if (H->d_part.reserve(H->part_len * sizeof(double)) != hipSuccess) return fail(H, "hipMalloc(partials) failed");
  H->hist_count = 0;
  H->last_resid = -1.0;
  if (H->cfg.extrapolate > 0) {
    if (H->d_hist.reserve((size_t)H->cfg.extrapolate * n * sizeof(double)) != hipSuccess ||
        H->d_img.reserve((size_t)H->cfg.extrapolate * n * sizeof(double)) != hipSuccess)
      return fail(H, "hipMalloc(history) failed");
    for (int k = 0; k < H->cfg.extrapolate; ++k) {
      H->hist[k] = H->d_hist.as<double>() + k * n;
      H->img[k] = H->d_img.as<double>() + k * n;
    }
  }
  if (!H->d_gdots.p &&
      H->d_gdots.alloc(GUESS_MAXV * (GUESS_MAXV + 1) / 2 * sizeof(double)) != hipSuccess)
    return fail(H, "hipMalloc(guess dots) failed");
  HIPCHECK(hipMemset(H->d_state.p, 0, sizeof(SolverState)));
  if (mixed_set_topology(H->mixed, H->nb, H->n, H->err) != 0) return -1;
  if (cnn_set_topology(H->cnn, H->nb, C.nbr.as<int>(), C.h.as<double>(), H->err) != 0)
    return -1;
  HIPCHECK(hipDeviceSynchronize());
  commit.ok = true;
  return 0;
}

static int solve_fp64(cup3d_poisson_t H, const double* d_rhs, double* d_x, const double* d_x0,
                      cup3d_solve_report* rep) {
  hipStream_t s = H->stream;
  SolverState init = make_state(H->cfg);
  *H->h_state = init;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  HIPCHECK(hipMemcpyAsync(st, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice, s));
  const long long zero_row = (H->cfg.bmean == 1 || H->cfg.bmean == 3) && H->topo.origin_block >= 0
                                 ? (long long)H->topo.origin_block * BLEN
                                 : -1;
  launch_getvec64(s, H->n, zero_row, d_rhs, d_x0, H->v64.b, H->v64.x[0], st, part, H->cfg.tol_rhs);
  if (run_bicgstab64(H, H->g64) != 0) return -1;
  launch_copy_opt64(s, H->n, H->v64, st, d_x);
  // true residual ||b - A x_opt|| in FP64 (into nu)
  launch_spmv64(s, SPMV_RESID, H->g64, d_x, nullptr, H->v64.b, H->v64.nu, nullptr, st, part);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(H->h_state, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  const SolverState& S = *H->h_state;
  if (rep) {
    rep->status = S.status;
    rep->iterations = S.iters_done;
    rep->restarts = S.restarts;
    rep->refinements = 1;
    rep->err_init = S.err_init;
    rep->err_final = S.err_opt;
    rep->rhs_norm = std::sqrt(S.bnorm2);
    rep->true_rel_resid = S.bnorm2 > 0 ? std::sqrt(S.true_rr / S.bnorm2) : std::sqrt(S.true_rr);
    rep->guess_resid = S.err_init;
  }
  return S.status;
}

// Increment form of a step: b and x0 as getVec makes them (v64.b, v64.x[0]),
// and r0 = b - A x0 in v64.nu; returns ||r0|| (FP64) or -1 on error.
static double increment_residual(cup3d_poisson_t H, const double* d_rhs, const double* d_prev) {
  hipStream_t s = H->stream;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  *H->h_state = make_state(H->cfg);
  if (hipMemcpyAsync(st, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(H, "hipMemcpyAsync(state) failed");
  const long long zero_row = (H->cfg.bmean == 1 || H->cfg.bmean == 3) && H->topo.origin_block >= 0
                                 ? (long long)H->topo.origin_block * BLEN
                                 : -1;
  launch_getvec64(s, H->n, zero_row, d_rhs, d_prev, H->v64.b, H->v64.x[0], st, part, 0.0);
  launch_spmv64(s, SPMV_RESID, H->g64, H->v64.x[0], nullptr, H->v64.b, H->v64.nu, nullptr, st,
                part);
  if (hipMemcpyAsync(H->h_state, st, sizeof(SolverState), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(H, "increment residual failed");
  return std::sqrt(H->h_state->true_rr);
}

// Solves the nk x nk system G g = f (symmetric positive semi-definite, nk
// <= GUESS_MAXV - 1) by Gaussian elimination with partial pivoting; false if
// it is singular to rounding.
static bool small_solve(int nk, double G[][GUESS_MAXV], const double* f, double* g) {
  double M[GUESS_MAXV][GUESS_MAXV + 1];
  double scale = 0;
  for (int i = 0; i < nk; ++i) {
    for (int j = 0; j < nk; ++j) M[i][j] = G[i][j];
    M[i][nk] = f[i];
    scale = std::max(scale, std::fabs(G[i][i]));
  }
  for (int c = 0; c < nk; ++c) {
    int p = c;
    for (int r = c + 1; r < nk; ++r)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    if (!(std::fabs(M[p][c]) > 1e-12 * scale)) return false;
    for (int j = 0; j <= nk; ++j) std::swap(M[c][j], M[p][j]);
    for (int r = c + 1; r < nk; ++r) {
      const double m = M[r][c] / M[c][c];
      for (int j = c; j <= nk; ++j) M[r][j] -= m * M[c][j];
    }
  }
  for (int c = nk - 1; c >= 0; --c) {
    double v = M[c][nk];
    for (int j = c + 1; j < nk; ++j) v -= M[c][j] * g[j];
    g[c] = v / M[c][c];
  }
  for (int c = 0; c < nk; ++c)
    if (!std::isfinite(g[c])) return false;
  return true;
}

// Weight of the history-noise estimate in the warm start's option score
// (see warm_start): a unit of residual carried in from the kept pressures'
// rounding counts as much as NOISE_WEIGHT units of the extrapolation's own
// residual. With 0 the deepest option wins whenever its residual is 1 %
// lower; on moving-blob sequences solved to 1e-7 and fed back (the AMR
// school, uniform 256x128x128 and AMR sphere grids) 120..300 selects the
// depths that need the fewest iterations (4, 5, 5 of 6).
constexpr double NOISE_WEIGHT = 200.0;

// Warm start next to the given guess x0 (the caller's pressure of the
// previous step), in the spirit of Fischer's projection onto previous
// solutions (Comput. Methods Appl. Mech. Engrg. 163, 1998), here with the
// residual norm: up to cfg.extrapolate directions from the differences of
// the previous pressures, d_1 = x0 - h_1, d_2 = h_1 - h_2, ..., and the
// learned net's step d_net = pred - x0 (CUP3D_GUESS_CNN, with trained
// weights). The guess becomes x0 + sum_k g_k d_k with the weights that
// minimise the FP64 residual ||r0 - sum_k g_k A d_k||, r0 = b - A x0 (least
// squares on the Gram matrix of [r0, A d_1, ...], options: the first m time
// directions, m = 0..M, with or without the net; a larger option replaces
// the best smaller one only if it lowers the squared residual by at least
// 1 %); g = 0 is one of the choices, so the result is never worse than x0,
// which is kept (the fallback) when no direction helps.
static int warm_start(cup3d_poisson_t H, const double* d_rhs, double* d_x, int* used,
                      double* resid, double* w_net, double* w_time, int* depth) {
  hipStream_t s = H->stream;
  *used = CUP3D_GUESS_GIVEN;
  *w_net = *w_time = 0.0;
  *depth = 0;
  // a failure on the way leaves history and images out of step: start over
  struct ClearUnlessDone {
    cup3d_poisson_t H;
    bool done = false;
    ~ClearUnlessDone() {
      if (!done) H->hist_count = 0;
    }
  } guard{H};
  const double r_prev = increment_residual(H, d_rhs, d_x);  // v64.b, v64.x[0], r0 in v64.nu
  if (r_prev < 0) return -1;
  *resid = r_prev;
  const KrylovVectors<double>& v = H->v64;
  const double* x0 = v.x[0];
  const long long n = H->n;
  // With bMean 1 or 3 getVec zeroes the origin cell of the guess (x0 holds
  // that zero); every direction is zero there too, so the weights minimise
  // the residual the solve actually starts from.
  const long long zero_row = (H->cfg.bmean == 1 || H->cfg.bmean == 3) && H->topo.origin_block >= 0
                                 ? (long long)H->topo.origin_block * BLEN
                                 : -1;
  // the images u_k = A d_k of the time directions, then the net's: d_k of
  // this solve is d_{k-1} of the last (h_k here is h_{k-1} there, bit for
  // bit), so A d_2 .. A d_m are the last solve's A d_1 .. A d_{m-1}, kept
  // in img[]; only A d_1 is new (into the slot of the dropped A d_K). Every
  // direction passes through `dir` on its way to the SpMV.
  double* const dir = v.t;
  GuessVecs gv{};
  gv.v[0] = v.nu;
  int nc = 0;
  const int K = H->cfg.extrapolate;
  const int mt = std::min(K, H->hist_count);  // time directions available
  if (K > 0) {
    double* fresh = H->img[K - 1];
    for (int k = K - 1; k > 0; --k) H->img[k] = H->img[k - 1];
    H->img[0] = fresh;
  }
  if (mt > 0) {
    launch_guess_direction(s, n, x0, H->hist[0], zero_row, dir);
    launch_spmv64(s, SPMV_PLAIN, H->g64, dir, H->img[0], nullptr, nullptr, nullptr,
                  H->d_state.as<SolverState>(), H->d_part.as<double>());
    if (hipGetLastError() != hipSuccess) return fail(H, "warm-start launch failed");
  }
  for (int k = 0; k < mt; ++k) gv.v[++nc] = H->img[k];
  int inet = -1;
  if (H->cfg.guess == CUP3D_GUESS_CNN && H->cnn.have_weights) {
    if (cnn_predict(H->cnn, s, v.nu, x0, H->cnn.d_pred, H->err) != 0) return -1;
    launch_guess_direction(s, n, H->cnn.d_pred, x0, zero_row, dir);
    launch_spmv64(s, SPMV_PLAIN, H->g64, dir, v.z2, nullptr, nullptr, nullptr,
                  H->d_state.as<SolverState>(), H->d_part.as<double>());
    gv.v[nc + 1] = v.z2;
    inet = nc++;
  }
  if (nc > 0) {
    constexpr int ND = GUESS_MAXV * (GUESS_MAXV + 1) / 2;
    double D[ND];
    launch_guess_gram(s, n, gv, nc + 1, H->d_part.as<double>(), H->d_gdots.as<double>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(D, H->d_gdots.p, sizeof(D), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    auto dot = [&](int a, int b) {  // index 0 = r0, k + 1 = u_k
      if (a > b) std::swap(a, b);
      return D[a * GUESS_MAXV - a * (a - 1) / 2 + (b - a)];
    };
    const double a2 = dot(0, 0);
    // The kept pressures carry the rounding of their own solves: residuals
    // rho_j of about the tolerance, whose error is the smooth part the
    // Krylov process removes slowest. An option x0 + sum_k g_k d_k takes
    // them with coefficients c_0 = 1 + g_1 (- g_net), c_j = g_{j+1} - g_j,
    // c_m = -g_m on x0, h_1 .. h_m (binomial for the extrapolations: their
    // norm grows like 2^m), so it carries their noise with a residual of
    // about sqrt(sum_j c_j^2 rho_j^2). Each option is scored by its
    // residual plus NOISE_WEIGHT times that estimate (in squares): deeper
    // extrapolation is taken only while what it removes exceeds the noise
    // it brings.
    const double rho0 = H->last_resid > 0 ? H->last_resid : 0.0;
    auto noise2 = [&](const double* g, int m, double gnet) {
      double c0 = 1.0 - gnet + (m > 0 ? g[0] : 0.0);
      double v = c0 * c0 * rho0 * rho0;
      for (int j = 1; j <= m; ++j) {
        const double cj = (j < m ? g[j] : 0.0) - g[j - 1];
        v += cj * cj * H->hist_rho[j - 1] * H->hist_rho[j - 1];
      }
      return v;
    };
    constexpr double NW2 = NOISE_WEIGHT * NOISE_WEIGHT;
    // options in order of size: the first m time directions, with or
    // without the net; each replaces the best so far only if it lowers the
    // score by at least 1 %
    constexpr double KEEP = 0.99;
    double best_r2 = a2, best_g[GUESS_MAXV] = {};
    double best_score = a2 + NW2 * rho0 * rho0;
    int best_m = 0;
    for (int size = 1; size <= nc; ++size) {
      for (int with_net = 0; with_net <= (inet >= 0 ? 1 : 0); ++with_net) {
        const int m = size - with_net;
        if (m < 0 || m > mt) continue;
        int idx[GUESS_MAXV], nk = 0;
        for (int k = 0; k < m; ++k) idx[nk++] = k;
        if (with_net) idx[nk++] = inet;
        double G[GUESS_MAXV][GUESS_MAXV], f[GUESS_MAXV], g[GUESS_MAXV];
        for (int i = 0; i < nk; ++i) {
          f[i] = dot(0, idx[i] + 1);
          for (int j = 0; j < nk; ++j) G[i][j] = dot(idx[i] + 1, idx[j] + 1);
        }
        if (!small_solve(nk, G, f, g)) continue;
        double r2 = a2;
        for (int i = 0; i < nk; ++i) r2 -= g[i] * f[i];
        const double score = r2 + NW2 * noise2(g, m, with_net ? g[nk - 1] : 0.0);
        if (!std::isfinite(score) || !(score < KEEP * best_score)) continue;
        best_score = score;
        best_r2 = r2;
        best_m = m;
        for (int k = 0; k < GUESS_MAXV; ++k) best_g[k] = 0;
        for (int i = 0; i < nk; ++i) best_g[idx[i]] = g[i];
      }
    }
    if (best_score < a2 + NW2 * rho0 * rho0) {
      const int m = best_m;  // time directions taken
      const double gnet = inet >= 0 ? best_g[inet] : 0.0;
      GuessVecs hv{};
      for (int k = 0; k < mt; ++k) hv.v[k] = H->hist[k];
      // written to d_x, not read by the solve before getVec copies it
      launch_guess_blend(s, n, x0, hv, best_g, m, H->cnn.d_pred, gnet, zero_row, d_x);
      HIPCHECK(hipGetLastError());
      *w_time = mt > 0 ? best_g[0] : 0.0;
      *w_net = gnet;
      *used = gnet != 0 ? CUP3D_GUESS_CNN : CUP3D_GUESS_GIVEN;
      *resid = std::sqrt(best_r2 > 0 ? best_r2 : 0.0);
      *depth = m;
    }
  }
  // this step's given guess (x0, as getVec made it) becomes h_1 of the next
  // solve; copied after the blend, which still reads the present history
  if (K > 0) {
    double* oldest = H->hist[K - 1];
    for (int k = K - 1; k > 0; --k) {
      H->hist[k] = H->hist[k - 1];
      H->hist_rho[k] = H->hist_rho[k - 1];
    }
    H->hist[0] = oldest;
    H->hist_rho[0] = H->last_resid > 0 ? H->last_resid : 0.0;
    HIPCHECK(hipMemcpyAsync(oldest, x0, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    H->hist_count = std::min(H->hist_count + 1, K);
  }
  guard.done = true;
  return 0;
}

int cup3d_poisson_solve_device(cup3d_poisson_t H, const double* d_rhs, double* d_x,
                               cup3d_solve_report* rep) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called before solve");
  if (check_vectors(H, {d_rhs, d_x}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  cup3d_solve_report local;
  std::memset(&local, 0, sizeof(local));
  HIPCHECK(hipEventRecord(H->ev0, H->stream));
  const double* x0 = d_x;
  int guess_used = H->cfg.guess;
  double guess_resid = -1;
  double w_net = 0, w_time = 0;
  int depth = 0;
  if (H->cfg.guess == CUP3D_GUESS_ZERO) x0 = nullptr;
  int st;
  if (H->cfg.guess == CUP3D_GUESS_CNN || (H->cfg.extrapolate && H->cfg.guess == CUP3D_GUESS_GIVEN)) {
    // learned and/or extrapolated warm start, the given guess as the fallback
    if (warm_start(H, d_rhs, d_x, &guess_used, &guess_resid, &w_net, &w_time, &depth) != 0)
      return -1;
  }
  if (H->cfg.precision == CUP3D_PREC_FP64)
    st = solve_fp64(H, d_rhs, d_x, x0, &local);
  else
    st = mixed_solve(H->mixed, H->stream, H->cfg, H->g64, H->g32, H->v64,
                     H->d_state.as<SolverState>(), H->h_state, H->d_part.as<double>(),
                     H->topo.origin_block, d_rhs, d_x, x0, &local, H->err);
  if (st < 0) return st;
  HIPCHECK(hipEventRecord(H->ev1, H->stream));
  HIPCHECK(hipEventSynchronize(H->ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, H->ev0, H->ev1));
  local.solve_ms = ms;
  local.guess_used = guess_used;
  if (guess_resid >= 0) local.guess_resid = guess_resid;
  local.guess_weight_net = w_net;
  local.guess_weight_time = w_time;
  local.guess_depth = depth;
  // what the next warm start takes as the noise of this solution
  const double r_true = local.true_rel_resid * local.rhs_norm;
  H->last_resid = std::isfinite(r_true) ? r_true : -1.0;
  if (rep) *rep = local;
  return st;
}

int cup3d_poisson_remove_mean(cup3d_poisson_t H, double* d_x, double* mean) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_x}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  double vol = 0.0;
  for (double h3 : H->topo.h3) vol += 512.0 * h3;
  const double inv_vol = 1.0 / vol;
  launch_remove_mean(H->stream, H->n, d_x, H->tp[H->tcur].h3.as<double>(), inv_vol,
                     H->d_part.as<double>(), H->d_gdots.as<double>());
  HIPCHECK(hipGetLastError());
  double wsum = 0.0;
  HIPCHECK(hipMemcpyAsync(&wsum, H->d_gdots.p, sizeof(double), hipMemcpyDeviceToHost, H->stream));
  HIPCHECK(hipStreamSynchronize(H->stream));
  if (mean) *mean = wsum * inv_vol;  // the value the kernel subtracted
  return 0;
}

int cup3d_poisson_solve(cup3d_poisson_t H, const double* const* rhs_blocks,
                        double* const* x_blocks, const double* rhs_contig, double* x_contig,
                        cup3d_solve_report* rep) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called before solve");
  if ((!rhs_blocks && !rhs_contig) || (!x_blocks && !x_contig)) return fail(H, "null vector");
  HIPCHECK(hipSetDevice(H->device));
  const size_t n = (size_t)H->n, bytes = n * sizeof(double);
  if (H->stage_n != H->n) {
    if (H->h_stage) (void)hipHostFree(H->h_stage);
    H->h_stage = nullptr;
    H->stage_n = 0;
    H->d_stage.release();
    HIPCHECK(hipHostMalloc((void**)&H->h_stage, 2 * bytes, hipHostMallocDefault));
    if (H->d_stage.alloc(2 * bytes) != hipSuccess) return fail(H, "hipMalloc(staging) failed");
    H->stage_n = H->n;
  }
  double *hb = H->h_stage, *hx = H->h_stage + n;
  double *db = H->d_stage.as<double>(), *dx = db + n;
  // getVec gather (PoissonSolverExp.cpp:389-417)
  parallel_blocks(H->nb, [&](int b0, int b1) {
    for (int b = b0; b < b1; ++b) {
      std::memcpy(hb + (size_t)b * BLEN,
                  rhs_blocks ? rhs_blocks[b] : rhs_contig + (size_t)b * BLEN, BLEN * sizeof(double));
      std::memcpy(hx + (size_t)b * BLEN, x_blocks ? x_blocks[b] : x_contig + (size_t)b * BLEN,
                  BLEN * sizeof(double));
    }
  });
  int st = -1;
  if (hipMemcpy(db, hb, bytes, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dx, hx, bytes, hipMemcpyHostToDevice) == hipSuccess) {
    st = cup3d_poisson_solve_device(H, db, dx, rep);
    if (st >= 0 && hipMemcpy(hx, dx, bytes, hipMemcpyDeviceToHost) == hipSuccess) {
      // scatter into the pressure blocks (.cpp:452-472)
      parallel_blocks(H->nb, [&](int b0, int b1) {
        for (int b = b0; b < b1; ++b)
          std::memcpy(x_blocks ? x_blocks[b] : x_contig + (size_t)b * BLEN, hx + (size_t)b * BLEN,
                      BLEN * sizeof(double));
      });
    } else if (st >= 0) {
      st = fail(H, "hipMemcpy D2H failed");
    }
  } else {
    st = fail(H, "hipMemcpy H2D failed");
  }
  return st;
}

int cup3d_poisson_apply_operator(cup3d_poisson_t H, const double* d_x, double* d_y) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_x, d_y}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  SolverState init = make_state(H->cfg);
  *H->h_state = init;
  HIPCHECK(hipMemcpyAsync(H->d_state.p, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice,
                          H->stream));
  launch_spmv64(H->stream, SPMV_PLAIN, H->g64, d_x, d_y, nullptr, nullptr, nullptr,
                H->d_state.as<SolverState>(), H->d_part.as<double>());
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_poisson_apply_preconditioner(cup3d_poisson_t H, const double* d_p, double* d_z) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_p, d_z}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  if (H->cfg.precision == CUP3D_PREC_FP64) {
    KrylovVectors<double> v = H->v64;
    v.p = const_cast<double*>(d_p);
    v.z1 = d_z;
    launch_precond_fp64(H, H->g64, PREC_PLAIN, v);
  } else {
    if (mixed_apply_precond(H->mixed, H->stream, H->cfg.precision, H->g32, d_p, d_z, H->err) != 0)
      return -1;
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_projection_rhs(cup3d_poisson_t H, const double* d_vel, const double* d_udef,
                         const double* d_chi, double dt, double* d_rhs) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_vel, d_rhs}) != 0) return -1;
  if ((d_udef == nullptr) != (d_chi == nullptr))
    return fail(H, "udef and chi must be given together (or both null)");
  if (d_udef && check_vectors(H, {d_udef, d_chi}) != 0) return -1;
  if (!(dt > 0) || !std::isfinite(dt)) return fail(H, "dt must be positive");
  if (d_rhs == d_vel || d_rhs == d_udef || d_rhs == d_chi)
    return fail(H, "rhs must not alias the inputs");
  HIPCHECK(hipSetDevice(H->device));
  launch_pressure_rhs(H->stream, H->g64, d_vel, d_udef, d_chi, dt, d_rhs);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_projection_increment(cup3d_poisson_t H, double* d_rhs, const double* d_p_old) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_rhs, d_p_old}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  // the plain operator (no bMean row): KernelDivPressure
  DevGrid<double> g = H->g64;
  g.bmean = 0;
  g.bmean_row = -1;
  SolverState init = make_state(H->cfg);
  *H->h_state = init;
  HIPCHECK(hipMemcpyAsync(H->d_state.p, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice,
                          H->stream));
  double* tmp = H->v64.nu;
  launch_spmv64(H->stream, SPMV_RESID, g, d_p_old, nullptr, d_rhs, tmp, nullptr,
                H->d_state.as<SolverState>(), H->d_part.as<double>());
  HIPCHECK(hipMemcpyAsync(d_rhs, tmp, (size_t)H->n * sizeof(double), hipMemcpyDeviceToDevice,
                          H->stream));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_projection_update_velocity(cup3d_poisson_t H, const double* d_p, double dt,
                                     double* d_vel) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_p, d_vel}) != 0) return -1;
  if (!(dt > 0) || !std::isfinite(dt)) return fail(H, "dt must be positive");
  HIPCHECK(hipSetDevice(H->device));
  launch_grad_p_update(H->stream, H->g64, d_p, dt, d_vel);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Implicit diffusion (AdvectionDiffusionImplicit.h:21-525,
// AdvectionDiffusionImplicit.cpp:328-430): the Helmholtz system of one
// velocity component on the same topology, FP64, classical BiCGSTAB with the
// DST block inverse of each level as the preconditioner.
// ---------------------------------------------------------------------------

// Domain faces whose velocity ghost is -u for component comp (BlockLabBC,
// Definitions.h:413-458): every wall face, and at freespace faces the
// component normal to the face.
static int diffusion_dmask(const cup3d_poisson_config& c, int comp) {
  int m = 0;
  for (int d = 0; d < 3; ++d)
    if (c.bc[d] == CUP3D_BC_WALL || (c.bc[d] == CUP3D_BC_FREESPACE && d == comp))
      m |= 3 << (2 * d);
  return m;
}

static int diffusion_grid(cup3d_poisson_t H, int comp, double nu, double dt, DevGrid<double>& g) {
  if (comp < 0 || comp > 2) return fail(H, "velocity component must be 0, 1 or 2");
  if (!(nu > 0) || !std::isfinite(nu) || !(dt > 0) || !std::isfinite(dt))
    return fail(H, "nu and dt must be positive and finite");
  const double sigma = 1.0 / (nu * dt);
  if (!std::isfinite(sigma)) return fail(H, "1 / (nu dt) is not finite");
  g = H->g64;
  g.bmean = 0;
  g.bmean_row = -1;
  g.helm = 1;
  g.dmask = diffusion_dmask(H->cfg, comp);
  g.helm_sigma = sigma;
  return 0;
}

// DiffusionSolver::solve on the device: b = d_rhs, x0 = d_x0 (null: 0);
// the best iterate is left in v64.p and its true residual in the report.
// Stopping rule of the reference (AdvectionDiffusionImplicit.h:235-236,
// 509): ||r|| < tol_abs or ||r|| / ||r0|| < tol_rel on the recurrence
// residual, at most cfg.max_iter iterations (1000 in the reference, :247),
// restarts unlimited (its cap is commented out, :487).
static int diffusion_solve(cup3d_poisson_t H, const DevGrid<double>& g, const double* d_rhs,
                           const double* d_x0, double tol_abs, double tol_rel,
                           cup3d_solve_report* rep) {
  hipStream_t s = H->stream;
  SolverState init = make_state(H->cfg);
  init.tol_abs = tol_abs;
  init.tol_rel = tol_rel;
  init.max_restarts = 1 << 30;
  *H->h_state = init;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  const KrylovVectors<double>& v = H->v64;
  HIPCHECK(hipEventRecord(H->ev0, s));
  HIPCHECK(hipMemcpyAsync(st, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice, s));
  launch_getvec64(s, H->n, -1, d_rhs, d_x0, v.b, v.x[0], st, part, 0.0);
  if (run_bicgstab64(H, g) != 0) return -1;
  launch_copy_opt64(s, H->n, v, st, v.p);
  launch_spmv64(s, SPMV_RESID, g, v.p, nullptr, v.b, v.nu, nullptr, st, part);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(H->h_state, st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipEventRecord(H->ev1, s));
  HIPCHECK(hipEventSynchronize(H->ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, H->ev0, H->ev1));
  const SolverState& S = *H->h_state;
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->status = S.status;
    rep->iterations = S.iters_done;
    rep->restarts = S.restarts;
    rep->refinements = 1;
    rep->err_init = S.err_init;
    rep->err_final = S.err_opt;
    rep->rhs_norm = std::sqrt(S.bnorm2);
    rep->true_rel_resid = S.bnorm2 > 0 ? std::sqrt(S.true_rr / S.bnorm2) : std::sqrt(S.true_rr);
    rep->solve_ms = ms;
    rep->guess_used = d_x0 ? CUP3D_GUESS_GIVEN : CUP3D_GUESS_ZERO;
    rep->guess_resid = S.err_init;
  }
  return S.status;
}

int cup3d_diffusion_solve_device(cup3d_poisson_t H, int comp, double nu, double dt,
                                 double tol_abs, double tol_rel, const double* d_rhs, double* d_x,
                                 cup3d_solve_report* rep) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called before solve");
  if (check_vectors(H, {d_rhs, d_x}) != 0) return -1;
  if (d_rhs == d_x) return fail(H, "rhs and x must not alias");
  DevGrid<double> g;
  if (diffusion_grid(H, comp, nu, dt, g) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  const int st = diffusion_solve(H, g, d_rhs, d_x, tol_abs, tol_rel, rep);
  if (st < 0) return st;
  HIPCHECK(hipMemcpyAsync(d_x, H->v64.p, (size_t)H->n * sizeof(double), hipMemcpyDeviceToDevice,
                          H->stream));
  HIPCHECK(hipStreamSynchronize(H->stream));
  return st;
}

int cup3d_diffusion_step_device(cup3d_poisson_t H, double nu, double dt, double tol_abs,
                                double tol_rel, const double* d_vel_old, double* d_vel,
                                cup3d_solve_report* reps) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called before solve");
  if (check_vectors(H, {d_vel_old, d_vel}) != 0) return -1;
  if (d_vel_old == d_vel) return fail(H, "vel_old and vel must not alias");
  HIPCHECK(hipSetDevice(H->device));
  hipStream_t s = H->stream;
  const KrylovVectors<double>& v = H->v64;
  int worst = 0;
  for (int c = 0; c < 3; ++c) {
    DevGrid<double> g;
    if (diffusion_grid(H, c, nu, dt, g) != 0) return -1;
    // rhs = h^3 (-(h lap u*)/h^3 + (u* - u^n)/(dt nu)) (KernelDiffusionRHS and
    // AdvectionDiffusionImplicit.cpp:376-400) = a - A u* with the Helmholtz
    // operator A = h lap - h^3 sigma and a = -h^3 sigma u^n: the residual
    // form of the operator, evaluated before component c is updated (the
    // Laplacian of component c reads component c only)
    *H->h_state = make_state(H->cfg);
    HIPCHECK(hipMemcpyAsync(H->d_state.p, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice,
                            s));
    launch_diffusion_gather(s, H->n, c, g.h3, g.helm_sigma, d_vel, d_vel_old, v.z1, v.z2);
    launch_spmv64(s, SPMV_RESID, g, v.z1, nullptr, v.z2, v.t, nullptr,
                  H->d_state.as<SolverState>(), H->d_part.as<double>());
    // x0 = 0 (the reference zeroes sim.pres before each solve, :398)
    const int st = diffusion_solve(H, g, v.t, nullptr, tol_abs, tol_rel, reps ? reps + c : nullptr);
    if (st < 0) return st;
    launch_diffusion_scatter(s, H->n, c, v.p, d_vel);
    HIPCHECK(hipGetLastError());
    worst = std::max(worst, st);
  }
  HIPCHECK(hipStreamSynchronize(s));
  return worst;
}

int cup3d_diffusion_apply_operator(cup3d_poisson_t H, int comp, double nu, double dt,
                                   const double* d_x, double* d_y) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_x, d_y}) != 0) return -1;
  DevGrid<double> g;
  if (diffusion_grid(H, comp, nu, dt, g) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  *H->h_state = make_state(H->cfg);
  HIPCHECK(hipMemcpyAsync(H->d_state.p, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice,
                          H->stream));
  launch_spmv64(H->stream, SPMV_PLAIN, g, d_x, d_y, nullptr, nullptr, nullptr,
                H->d_state.as<SolverState>(), H->d_part.as<double>());
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_diffusion_apply_preconditioner(cup3d_poisson_t H, double nu, double dt,
                                         const double* d_p, double* d_z) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_p, d_z}) != 0) return -1;
  DevGrid<double> g;
  if (diffusion_grid(H, 0, nu, dt, g) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  KrylovVectors<double> v = H->v64;
  v.p = const_cast<double*>(d_p);
  v.z1 = d_z;
  launch_precond_dst(H->stream, PREC_PLAIN, H->nb, H->d_dst.as<double>(), g.h, g.helm_sigma, v,
                     H->d_state.as<SolverState>());
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_cnn_num_weights(void) { return cnn_num_weights(); }

int cup3d_poisson_cnn_set_weights(cup3d_poisson_t H, const float* w, int nw) {
  if (!H) return -1;
  HIPCHECK(hipSetDevice(H->device));
  return cnn_set_weights(H->cnn, w, nw, H->err);
}

int cup3d_poisson_cnn_get_weights(cup3d_poisson_t H, float* w, int nw) {
  if (!H) return -1;
  HIPCHECK(hipSetDevice(H->device));
  return cnn_get_weights(H->cnn, w, nw, H->err);
}

int cup3d_poisson_cnn_predict(cup3d_poisson_t H, const double* d_rhs, const double* d_prev,
                              double* d_out) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_rhs, d_prev, d_out}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  if (increment_residual(H, d_rhs, d_prev) < 0) return -1;
  return cnn_predict(H->cnn, H->stream, H->v64.nu, H->v64.x[0], d_out, H->err);
}

int cup3d_poisson_cnn_add_sample(cup3d_poisson_t H, const double* d_rhs, const double* d_prev,
                                 const double* d_p) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_rhs, d_prev, d_p}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  if (increment_residual(H, d_rhs, d_prev) < 0) return -1;
  return cnn_add_sample(H->cnn, H->stream, H->v64.nu, H->v64.x[0], d_p, H->err);
}

int cup3d_poisson_cnn_clear_samples(cup3d_poisson_t H) {
  if (!H) return -1;
  HIPCHECK(hipSetDevice(H->device));
  cnn_clear_samples(H->cnn);
  return 0;
}

int cup3d_poisson_cnn_num_samples(cup3d_poisson_t H) {
  return H ? (int)H->cnn.samples.size() : -1;
}

int cup3d_poisson_cnn_train(cup3d_poisson_t H, int steps, int batch, float lr, float res_weight,
                            uint64_t seed, double* loss) {
  if (!H) return -1;
  HIPCHECK(hipSetDevice(H->device));
  return cnn_train(H->cnn, H->stream, steps, batch, lr, res_weight, seed, loss, H->err);
}

int cup3d_poisson_cnn_eval(cup3d_poisson_t H, int k, double* loss) {
  if (!H || !loss) return -1;
  HIPCHECK(hipSetDevice(H->device));
  return cnn_eval(H->cnn, H->stream, k, loss, H->err);
}

int cup3d_debug_cnn_gradient(cup3d_poisson_t H, int k, float res_weight, float* grad,
                             double* loss) {
  if (!H || !grad || !loss) return -1;
  HIPCHECK(hipSetDevice(H->device));
  return cnn_gradient(H->cnn, H->stream, k, res_weight, grad, loss, H->err);
}

int cup3d_debug_device_topology(cup3d_poisson_t H, int32_t* nbr, int8_t* kind, int32_t* fnb,
                                int8_t* bpar, double* hh, double* invh, double* h3,
                                int64_t* info) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  HIPCHECK(hipSetDevice(H->device));
  const DevTopo& C = H->tp[H->tcur];
  const size_t nb = (size_t)H->nb;
  auto get = [&](void* dst, const DevBuf& src, size_t bytes) -> hipError_t {
    return dst ? hipMemcpy(dst, src.p, bytes, hipMemcpyDeviceToHost) : hipSuccess;
  };
  HIPCHECK(get(nbr, C.nbr, nb * 6 * sizeof(int32_t)));
  HIPCHECK(get(kind, C.fkind, nb * 6));
  HIPCHECK(get(fnb, C.fnb, nb * 24 * sizeof(int32_t)));
  HIPCHECK(get(bpar, C.bpar, nb));
  HIPCHECK(get(hh, C.h, nb * sizeof(double)));
  HIPCHECK(get(invh, C.invh, nb * sizeof(double)));
  HIPCHECK(get(h3, C.h3, nb * sizeof(double)));
  if (info) {
    info[0] = H->topo.origin_block;
    info[1] = H->topo.n_coarse_faces;
    info[2] = H->topo.n_fine_faces;
  }
  return 0;
}

int cup3d_debug_apply_operator32(cup3d_poisson_t H, const double* d_x, double* d_y) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (check_vectors(H, {d_x, d_y}) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  if (mixed_apply_operator32(H->mixed, H->stream, H->g32, H->d_state.as<SolverState>(),
                             H->d_part.as<double>(), d_x, d_y, H->err) != 0)
    return -1;
  HIPCHECK(hipStreamSynchronize(H->stream));
  return 0;
}

int cup3d_debug_time_kernel(cup3d_poisson_t H, int which, int reps, double* avg_ms,
                            double* bytes, double* flops) {
  if (!H) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (reps <= 0 || which < 0 || which > 4) return fail(H, "bad arguments");
  HIPCHECK(hipSetDevice(H->device));
  hipStream_t s = H->stream;
  SolverState init = make_state(H->cfg);
  init.tol_abs = -1.0;  // the update kernel never declares convergence here
  init.tol_rel = -1.0;
  init.max_iter = 0x7fffffff;
  init.max_restarts = 0;
  init.err_init = 1.0;
  init.iter = 1;
  init.beta = 0.0;      // K1 then recomputes p = r on every repetition
  init.p_zero = 0;
  *H->h_state = init;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  HIPCHECK(hipMemcpyAsync(st, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice, s));
  const double N = (double)H->n;
  double by = 0, fl = 0;
  if (H->cfg.precision != CUP3D_PREC_FP64) {
    H->mixed.timing_precision = H->cfg.precision;
    if (mixed_time_kernel(H->mixed, s, which, reps, H->g32, st, part, H->ev0, H->ev1, avg_ms, &by,
                          &fl, H->err) != 0)
      return -1;
  } else {
    const KrylovVectors<double>& v = H->v64;
    auto launch = [&]() {
      switch (which) {
        case 0: launch_spmv64(s, SPMV_PLAIN, H->g64, v.z1, v.nu, nullptr, nullptr, nullptr, st, part); break;
        case 1: launch_precond_fp64(H, H->g64, PREC_PLAIN, v); break;
        case 2: launch_update64(s, H->n, v, st, part); break;
        case 3: launch_precond_fp64(H, H->g64, PREC_P, v); break;
        default: launch_spmv64(s, SPMV_DOT1, H->g64, v.z1, v.nu, v.rhat, nullptr, nullptr, st, part); break;
      }
    };
    launch();  // warm
    HIPCHECK(hipEventRecord(H->ev0, s));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHECK(hipEventRecord(H->ev1, s));
    HIPCHECK(hipEventSynchronize(H->ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, H->ev0, H->ev1));
    *avg_ms = ms / reps;
    const double words[5] = {2, 2, 8, 5, 3};
    const bool dst = H->cfg.precond == CUP3D_PRECOND_DST;
    const double fpc[5] = {8, dst ? 96.0 : 1024.0, 8, dst ? 96.0 + 4 : 1024.0 + 4, 10};
    by = 8.0 * words[which] * N;
    fl = fpc[which] * N;
  }
  if (bytes) *bytes = by;
  if (flops) *flops = fl;
  return 0;
}

int cup3d_debug_time_diffusion_kernel(cup3d_poisson_t H, int which, double nu, double dt,
                                      int reps, double* avg_ms, double* bytes, double* flops) {
  if (!H || !avg_ms) return -1;
  if (H->nb <= 0) return fail(H, "set_topology must be called first");
  if (reps <= 0 || which < 0 || which > 2) return fail(H, "bad arguments");
  DevGrid<double> g;
  if (diffusion_grid(H, 0, nu, dt, g) != 0) return -1;
  HIPCHECK(hipSetDevice(H->device));
  hipStream_t s = H->stream;
  SolverState init = make_state(H->cfg);
  init.beta = 0.0;  // K1 recomputes p = r on every repetition
  init.iter = 1;
  *H->h_state = init;
  SolverState* st = H->d_state.as<SolverState>();
  double* part = H->d_part.as<double>();
  HIPCHECK(hipMemcpyAsync(st, H->h_state, sizeof(SolverState), hipMemcpyHostToDevice, s));
  const KrylovVectors<double>& v = H->v64;
  HIPCHECK(hipMemsetAsync(v.r, 0, (size_t)H->n * sizeof(double), s));
  HIPCHECK(hipMemsetAsync(v.nu, 0, (size_t)H->n * sizeof(double), s));
  HIPCHECK(hipMemsetAsync(v.p, 0, (size_t)H->n * sizeof(double), s));
  HIPCHECK(hipMemsetAsync(v.z1, 0, (size_t)H->n * sizeof(double), s));
  auto launch = [&]() {
    if (which == 0)
      launch_spmv64(s, SPMV_PLAIN, g, v.z1, v.nu, nullptr, nullptr, nullptr, st, part);
    else
      launch_precond_dst(s, which == 1 ? PREC_P : PREC_PLAIN, H->nb, H->d_dst.as<double>(), g.h,
                         g.helm_sigma, v, st);
  };
  launch();  // warm
  HIPCHECK(hipEventRecord(H->ev0, s));
  for (int r = 0; r < reps; ++r) launch();
  HIPCHECK(hipEventRecord(H->ev1, s));
  HIPCHECK(hipEventSynchronize(H->ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, H->ev0, H->ev1));
  *avg_ms = ms / reps;
  const double N = (double)H->n;
  const double words[3] = {2, 5, 2};
  const double fpc[3] = {10, 96 + 5, 96};
  if (bytes) *bytes = 8.0 * words[which] * N;
  if (flops) *flops = fpc[which] * N;
  return 0;
}

int64_t cup3d_poisson_num_coarse_fine_faces(cup3d_poisson_t H) {
  return H ? (int64_t)H->topo.n_coarse_faces + H->topo.n_fine_faces : -1;
}

void cup3d_poisson_destroy(cup3d_poisson_t H) {
  if (!H) return;
  (void)hipSetDevice(H->device);
  (void)hipStreamSynchronize(H->stream);
  cnn_free(H->cnn);
  mixed_free(H->mixed);
  if (H->h_state) (void)hipHostFree(H->h_state);
  if (H->h_stage) (void)hipHostFree(H->h_stage);
  if (H->ev0) (void)hipEventDestroy(H->ev0);
  if (H->ev1) (void)hipEventDestroy(H->ev1);
  if (H->stream) (void)hipStreamDestroy(H->stream);
  delete H;
}

double cup3d_debug_copy_rate(int device, int64_t bytes, int reps) {
  if (bytes <= 0 || bytes % 16 || reps <= 0) return -1.0;
  if (hipSetDevice(device) != hipSuccess) return -1.0;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double rate = -1.0;
  if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
      hipMemset(a, 0, bytes) == hipSuccess) {
    launch_stream_copy(nullptr, bytes / 16, a, b);  // warm
    float ms = 0.f;
    if (hipEventRecord(e0, nullptr) == hipSuccess) {
      for (int r = 0; r < reps; ++r) launch_stream_copy(nullptr, bytes / 16, a, b);
      if (hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
          hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f)
        rate = 2.0 * (double)bytes * reps / (ms * 1e-3);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rate;
}

int cup3d_debug_profile(cup3d_poisson_t H, int enable) {
  if (!H) return -1;
  if (H->cfg.precision == CUP3D_PREC_FP64) return fail(H, "in-solve profiling covers the mixed form");
  HIPCHECK(hipSetDevice(H->device));
  MixedWorkspace& M = H->mixed;
  if (enable && !M.pev[0])
    for (hipEvent_t& e : M.pev) HIPCHECK(hipEventCreate(&e));
  M.prof = enable != 0;
  for (double& m : M.prof_ms) m = 0.0;
  M.prof_n = 0;
  return 0;
}

int cup3d_debug_profile_read(cup3d_poisson_t H, double* avg_us, long long* samples) {
  if (!H || !avg_us || !samples) return -1;
  MixedWorkspace& M = H->mixed;
  for (int k = 0; k < 5; ++k) avg_us[k] = M.prof_n ? 1e3 * M.prof_ms[k] / M.prof_n : 0.0;
  *samples = M.prof_n;
  for (double& m : M.prof_ms) m = 0.0;
  M.prof_n = 0;
  return 0;
}

const char* cup3d_poisson_last_error(cup3d_poisson_t H) {
  return H ? H->err.c_str() : g_create_error.c_str();
}

}  // extern "C"
