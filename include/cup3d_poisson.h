/*
 * cup3d_poisson.h — C ABI of the MI355X pressure-Poisson solver for
 * CubismUP_3D (libcup3d_poisson.so).
 *
 * Drop-in boundary. The reference selects its solver with
 *   std::shared_ptr<PoissonSolverBase> makePoissonSolver(SimulationData&)
 *     (source/poisson/PoissonSolverBase.cpp:10-32)
 * and calls the argument-less
 *   virtual void PoissonSolverBase::solve()
 *     (source/poisson/PoissonSolverBase.h:10-17)
 * once per time step from PressureProjection (PressureProjection.cpp:402).
 * solve() reads the rhs from the blocks of sim.lhs and the initial guess from
 * sim.pres, and writes the pressure back into sim.pres
 * (PoissonSolverExp.cpp:375-418, 452-472). The functions below carry exactly
 * that data across a plain C boundary: no C++ types, no exceptions, plain
 * pointers and sizes. A C++ adapter implementing PoissonSolverBase on top of
 * them is cup3d_amd/host/cup3d_poisson_adapter.hpp.
 *
 * Layout. A block is 8x8x8 doubles stored [iz][iy][ix] (Cubism ScalarBlock,
 * source/Definitions.h:524-531). Blocks are numbered by the caller (Cubism's
 * blockID); the flat cell index is blockID*512 + iz*64 + iy*8 + ix
 * (PoissonSolverExp.h:69-70).
 *
 * Errors. Every function returns an int: >= 0 success (for solves: the
 * status below), < 0 failure; cup3d_poisson_last_error() gives the message.
 * Non-convergence is not an error: like the reference (BiCGSTAB.cu:612-625)
 * the best iterate is returned with status CUP3D_MAXITER.
 *
 * Threading. One host thread per handle at a time; each handle owns one HIP
 * stream on its device; calls block until the result is written.
 */
#ifndef CUP3D_POISSON_H
#define CUP3D_POISSON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cup3d_poisson* cup3d_poisson_t;

/* BCflag, source/Definitions.h:38 */
enum { CUP3D_BC_FREESPACE = 0, CUP3D_BC_PERIODIC = 1, CUP3D_BC_WALL = 2 };

/* Solve status (return value of the solve calls). */
enum {
  CUP3D_CONVERGED = 0,  /* ||r|| <= tol_abs or ||r||/||r0|| <= tol_rel (BiCGSTAB.cu:595) */
  CUP3D_MAXITER = 1,    /* iteration cap reached; best iterate returned (:612-625) */
  CUP3D_BREAKDOWN = 2,  /* restarts exhausted (:496-499); best iterate returned */
  CUP3D_NONFINITE = 3   /* residual became NaN/Inf; best iterate returned */
};

/* Arithmetic of the solve. */
enum {
  CUP3D_PREC_FP64 = 0,        /* drop-in: FP64 throughout */
  CUP3D_PREC_MIXED_FP16 = 1,  /* FP16 P_inv MFMA, FP32 Krylov, FP64 refinement */
  CUP3D_PREC_MIXED_BF16 = 2   /* BF16 P_inv MFMA, FP32 Krylov, FP64 refinement */
};

/* Initial guess. */
enum {
  CUP3D_GUESS_GIVEN = 0,  /* x on input is the guess (sim.pres, the reference contract) */
  CUP3D_GUESS_ZERO = 1,   /* x on input ignored, guess = 0 */
  CUP3D_GUESS_CNN = 2     /* learned warm start from (rhs, x on input = previous pressure);
                             falls back to the given x when its residual is larger */
};
/* With cfg.extrapolate = K in 1..8 (and guess GIVEN or CNN) the solver keeps
 * the x handed to its last K solves on the same topology (the caller's
 * pressures of the previous steps, h_1 the most recent) and tries the
 * directions d_1 = x - h_1, d_2 = h_1 - h_2, ..., d_K = h_{K-1} - h_K (those
 * the history holds). The guess used is x + sum_k g_k d_k + g_net (net - x)
 * with the weights that minimise the FP64 residual ||b - A guess|| (least
 * squares over r0 = b - A x and the directions' images A d_k: projection onto
 * the recent solutions, as in Fischer, CMAME 163 (1998) 193-204, in the
 * residual norm). K = 1 with g_1 = 1 is the linear extrapolation in time
 * 2 x - h_1; K = 2, 3, ... reach the quadratic, cubic, ... extrapolations.
 * The time directions are taken nested (d_1, then d_2, ...) and the net's as
 * an option beside them. Each option is scored by its squared residual plus
 * 200^2 times the rounding it carries in from the kept pressures,
 * sum_j c_j^2 rho_j^2 (c_j its coefficients on x, h_1, ...; rho_j the true
 * residual of the solve each came from, assuming the caller hands the
 * solutions back), which stops the extrapolation where its binomial weights
 * would amplify that rounding; a larger set is only taken if it lowers the
 * score by at least 1 % against the best smaller one, and the given x (all
 * weights 0) is one of the choices, so the guess's residual is never above
 * x's. guess_weight_time reports g_1, guess_depth the number of time
 * directions taken. set_topology clears the history. Cost per
 * solve: one FP64 operator application (A d_2 .. A d_K are the last solve's
 * A d_1 .. A d_{K-1}, kept), one Gram reduction, one blend; memory: 2 K
 * vectors of n doubles. */

typedef struct {
  int bpd[3];        /* blocks per direction at level 0 (-bpdx/-bpdy/-bpdz) */
  int level_max;     /* -levelMax (informational: any level present in the block list is accepted) */
  int bc[3];         /* BCflag per axis (-BC_x/-BC_y/-BC_z); freespace and wall are zero-flux */
  double h0;         /* cell size at level 0 */
  double tol_abs;    /* sim.PoissonErrorTol (-poissonTol), absolute on ||r||_2 */
  double tol_rel;    /* sim.PoissonErrorTolRel (-poissonTolRel), relative to ||r0||_2 */
  double tol_rhs;    /* extra absolute tolerance tol_rhs*||b||_2 (0 = off); the
                        true-relative-residual stopping test used for benchmarks */
  int bmean;         /* sim.bMeanConstraint (-bMeanConstraint): 0, 1, 2 or 3 */
  int max_iter;      /* 1000 in the reference (BiCGSTAB.cu:474) */
  int max_restarts;  /* 100 in the reference (PoissonSolverExp.cpp:435) */
  int precision;     /* CUP3D_PREC_* */
  int guess;         /* CUP3D_GUESS_* */
  int device;        /* HIP device ordinal */
  /* mixed precision only: floor of the FP32 Krylov tolerance relative to the
     residual it starts from (it targets max(inner_tol, 0.5 tol / |r|);
     0 = 1e-7), and the cap on restarts of the FP32 Krylov process (each
     breakdown, iteration cap or stall starts a new one from the FP64
     residual; a stall is 50 iterations without halving the best residual,
     a window that doubles after every stall restart that did not halve the
     FP64 residual) */
  double inner_tol;
  int max_refine;
  int poll_every;    /* iterations between host polls of the device status (0 = default) */
  int gemm_passes;   /* mixed only: 0 (or 2, 3) = input split into hi + lo, two MFMA
                        passes, linear to FP32 precision; 1 = one FP16/BF16 pass */
  int extrapolate;   /* 0 = off (the reference contract); K = 1..8 = also extrapolate
                        the guess in time from the last K solves' (see above) */
  int precond;       /* how M^-1 = blockdiag(h_b^-1 P_inv) is applied.
                        CUP3D_PRECOND_GEMM (0): P_inv x block matrix on the matrix cores
                        (FP64, as the reference, BiCGSTAB.cu:526-528; the mixed forms:
                        FP16/BF16 with FP32 accumulate); CUP3D_PRECOND_DST (1): the same
                        exact inverse by 8-point sine transforms (P_inv is diagonalised
                        by the 3-D DST) in the form's own precision (FP64, or FP32 in the
                        mixed forms' inner solve), bound by HBM, not the matrix cores. */
} cup3d_poisson_config;

enum { CUP3D_PRECOND_GEMM = 0, CUP3D_PRECOND_DST = 1 };

typedef struct {
  int level;     /* BlockInfo::level */
  int index[3];  /* BlockInfo::index */
} cup3d_block_desc;

typedef struct {
  int status;            /* CUP3D_CONVERGED ... */
  int iterations;        /* Krylov iterations (sum over refinement steps) */
  int restarts;
  int refinements;       /* mixed: FP64 corrections (fold of the FP32 increment into
                            the FP64 iterate + true residual); 1 for FP64 */
  double err_init;       /* ||b - A x0||_2 (BiCGSTAB.cu:461) */
  double err_final;      /* best recurrence residual ||r||_2 */
  double rhs_norm;       /* ||b||_2 */
  double true_rel_resid; /* ||b - A x||_2 / ||b||_2 recomputed in FP64 */
  double solve_ms;       /* device time of the solve (events around it) */
  int guess_used;        /* CUP3D_GUESS_* actually used (after CNN fallback) */
  double guess_resid;    /* ||b - A x0||_2 of the guess used */
  double guess_weight_net;   /* g_net of the warm start (0: the net's guess not used) */
  double guess_weight_time;  /* g_time of the warm start (0: no extrapolation) */
  int guess_depth;           /* time directions of the warm start taken (0: none) */
} cup3d_solve_report;

/* Create a solver on cfg->device. Returns 0 or < 0. */
int cup3d_poisson_create(const cup3d_poisson_config* cfg, cup3d_poisson_t* out);

/* Describe the grid: nblocks leaves in blockID order. Builds neighbour and
 * coarse-fine interface tables and uploads them. Call again after every mesh
 * change (the reference rebuilds its matrix when sim.pres->UpdateFluxCorrection
 * is set, PoissonSolverExp.cpp:437-442). */
int cup3d_poisson_set_topology(cup3d_poisson_t h, int64_t nblocks, const cup3d_block_desc* blocks);

/* Host solve over Cubism block pointers: rhs_blocks[b] and x_blocks[b] each
 * point at 512 doubles. If the pointer arrays are NULL, rhs and x are single
 * contiguous [nblocks][512] arrays passed as rhs_contig / x_contig. */
int cup3d_poisson_solve(cup3d_poisson_t h, const double* const* rhs_blocks,
                        double* const* x_blocks, const double* rhs_contig, double* x_contig,
                        cup3d_solve_report* rep);

/* Device solve: d_rhs and d_x are device pointers to [nblocks][512] doubles
 * (HBM-resident, on the handle's device). d_x holds the guess on input. */
int cup3d_poisson_solve_device(cup3d_poisson_t h, const double* d_rhs, double* d_x,
                               cup3d_solve_report* rep);

/* The caller's projection after a solve (PressureProjection.cpp:404-431) on
 * the device: x -= sum_b h_b^3 sum_cells x / sum_b 512 h_b^3 (the
 * volume-weighted mean, summed in a fixed order). The mean removed is
 * returned in *mean when mean is not NULL. */
int cup3d_poisson_remove_mean(cup3d_poisson_t h, double* d_x, double* mean);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). cup3d_poisson_remove_mean without a handle and without
 * any allocation: d_h3 holds h_b^3 per block (nblocks doubles), d_x the
 * field ([nblocks][512], 16-byte aligned), both on the current device. The
 * partial sums of the reduction live in a static device array of the
 * library sized for CUP3D_NOALLOC_MAX_BLOCKS blocks (4 MiB per GPU, held
 * until the process ends); larger inputs are rejected on the host. Launches
 * on `stream` (a hipStream_t; NULL = the default stream). If mean is not
 * NULL the removed mean is copied to *mean and the call waits for the
 * stream; with mean NULL it only enqueues (capturable). The array is shared
 * by every call on a GPU: calls that overlap on one GPU race. Results are
 * bitwise those of cup3d_poisson_remove_mean on the same grid. */
#define CUP3D_NOALLOC_MAX_BLOCKS (1 << 20)
int cup3d_remove_mean_noalloc(int64_t nblocks, const double* d_h3, double* d_x, void* stream,
                              double* mean);

/* The caller around the solve, on the handle's grid and stream
 * (PressureProjection::operator(), PressureProjection.cpp:333-466). Vector
 * fields are Cubism VectorBlocks: [nblocks][8][8][8][3] doubles ([iz][iy][ix],
 * then the component); scalars as the solver's vectors. At coarse-fine faces
 * the finer side takes its ghosts from the solver's interface interpolation
 * (exact for quadratics) and the coarser side receives the four fine face
 * terms (Cubism's flux correction); at non-periodic domain faces the normal
 * velocity is mirrored (BlockLabBC: walls and freespace) and the pressure
 * has zero gradient (BlockLabNeumann3D). */
/* KernelPressureRHS (PressureProjection.cpp:96-199): d_rhs = h^2/(2 dt) div(vel)
 * - chi h^2/(2 dt) div(udef) by central differences; d_udef and d_chi are an
 * obstacle's deformation velocity and indicator, both given or both NULL. */
int cup3d_projection_rhs(cup3d_poisson_t h, const double* d_vel, const double* d_udef,
                         const double* d_chi, double dt, double* d_rhs);
/* The increment form of the solve (PressureProjection.cpp:340-362 with
 * KernelDivPressure :16-94): d_rhs -= A d_p_old with the plain operator (no
 * bMean row); the caller then solves from a zero guess and adds d_p_old back
 * after removing the mean. */
int cup3d_projection_increment(cup3d_poisson_t h, double* d_rhs, const double* d_p_old);
/* KernelGradP and the velocity update (PressureProjection.cpp:244-323, 449-466):
 * vel += -(dt h^2 / 2) (p(x+h) - p(x-h)) / h^3 per component, in place. */
int cup3d_projection_update_velocity(cup3d_poisson_t h, const double* d_p, double dt,
                                     double* d_vel);

/* y = A x with the solver's operator and its bMean handling: the bMean = 1
 * row, the bMean = 2 term (sum h^3 x) h^3, the bMean = 3 row (device
 * pointers). */
int cup3d_poisson_apply_operator(cup3d_poisson_t h, const double* d_x, double* d_y);

/* z = blockdiag(h_b^-1 P_inv) p on the device, in the handle's precision. */
int cup3d_poisson_apply_preconditioner(cup3d_poisson_t h, const double* d_p, double* d_z);

/* Implicit diffusion (-implicitDiffusion; AdvectionDiffusionImplicit.h:21-525,
 * AdvectionDiffusionImplicit.cpp:328-430) on the handle's grid, in FP64
 * whatever the handle's precision. The system of velocity component comp
 * (0, 1, 2) is KernelLHSDiffusion's (AdvectionDiffusionImplicit.h:40-56):
 *   A x = h (sum of the six lab neighbours - 6 x) - h^3 x / (nu dt)
 * with the solver's coarse-fine interface and, at non-periodic domain faces,
 * BlockLabBC's velocity ghosts (Definitions.h:58-122, 287-330, 413-458):
 * -x on walls, and on freespace faces -x for the normal component and +x for
 * the tangential ones. Classical BiCGSTAB (the state machine of the pressure
 * solve) preconditioned by the exact block inverse h^-1 A_loc^-1 of each
 * block's level, applied by discrete sine transforms (the reference
 * approximates the same inverse by per-block CG,
 * DiffusionSolverAMRKernels.cpp:149-213). Stopping rule of DiffusionSolver::
 * solve: ||r|| < tol_abs or ||r|| / ||r0|| < tol_rel (sim.DiffusionErrorTol,
 * default 1e-6, and sim.DiffusionErrorTolRel, default 1e-4,
 * SimulationData.cpp:92-93), at most cfg.max_iter iterations, restarts
 * unlimited. Returns the status like the pressure solve. */
/* DiffusionSolver::solve for one component: d_x holds the guess on input
 * (the reference passes 0) and receives the best iterate. */
int cup3d_diffusion_solve_device(cup3d_poisson_t h, int comp, double nu, double dt,
                                 double tol_abs, double tol_rel, const double* d_rhs, double* d_x,
                                 cup3d_solve_report* rep);
/* AdvectionDiffusionImplicit::euler after its advection kernel
 * (AdvectionDiffusionImplicit.cpp:354-419): d_vel_old = u^n and d_vel = u*
 * (the velocity after the explicit advection), both [nblocks][8][8][8][3];
 * for each component c: rhs = -h lap(u*_c) + h^3 (u*_c - u^n_c) / (nu dt)
 * (KernelDiffusionRHS, :19-125, and :376-400), x0 = 0, solve, u*_c += x.
 * d_vel holds u^{n+1} on return; reps (3 reports, or NULL) one per component.
 * Returns the worst status of the three solves. */
int cup3d_diffusion_step_device(cup3d_poisson_t h, double nu, double dt, double tol_abs,
                                double tol_rel, const double* d_vel_old, double* d_vel,
                                cup3d_solve_report* reps);
/* y = A x of component comp (DiffusionSolver::_lhs, AdvectionDiffusionImplicit.h:151-190). */
int cup3d_diffusion_apply_operator(cup3d_poisson_t h, int comp, double nu, double dt,
                                   const double* d_x, double* d_y);
/* z = h_b^-1 A_loc(h_b^2 / (nu dt))^-1 p per block (DiffusionSolver::
 * _preconditioner, AdvectionDiffusionImplicit.h:111-149, solved exactly). */
int cup3d_diffusion_apply_preconditioner(cup3d_poisson_t h, double nu, double dt,
                                         const double* d_p, double* d_z);

/* Learned warm start (precision-independent). A 3-layer 3-D CNN per 8^3
 * block with a one-cell halo predicts the pressure from (rhs, previous
 * pressure); inference runs on FP16 matrix cores, training in FP32, both by
 * this library's kernels. Training triples are copied into a device-resident
 * training set owned by the handle.
 * n_weights = cup3d_cnn_num_weights(). */
int cup3d_cnn_num_weights(void);
int cup3d_poisson_cnn_set_weights(cup3d_poisson_t h, const float* w, int n_weights);
int cup3d_poisson_cnn_get_weights(cup3d_poisson_t h, float* w, int n_weights);
/* d_out = p_prev + s_r * net(res / (h s_r), p_prev / s_p) on all blocks, with
 * res = b - A p_prev the residual of the previous pressure for this rhs
 * (b, p_prev as getVec makes them from d_rhs, d_prev), s_r = RMS(res / h),
 * s_p = RMS(p_prev). Device pointers, doubles. */
int cup3d_poisson_cnn_predict(cup3d_poisson_t h, const double* d_rhs, const double* d_prev,
                              double* d_out);
/* Append one triple (rhs, p_prev, p) (device pointers) to the training set;
 * stored as (res / h, p_prev, (p - p_prev) / s_r) in FP32. */
int cup3d_poisson_cnn_add_sample(cup3d_poisson_t h, const double* d_rhs, const double* d_prev,
                                 const double* d_p);
int cup3d_poisson_cnn_clear_samples(cup3d_poisson_t h);
int cup3d_poisson_cnn_num_samples(cup3d_poisson_t h);
/* `steps` Adam steps on minibatches of `batch` random (sample, block) items
 * drawn with `seed`. Loss in normalised units:
 *   0.5 mean (net - target)^2 + res_weight * 0.5 mean rho^2,
 * rho = 7-point Laplacian of the net's increment minus res / (h s_r) (the
 * local residual of the increment equation; the target closes the stencil
 * outside the block). loss[0] = first minibatch loss, loss[1] = last. */
int cup3d_poisson_cnn_train(cup3d_poisson_t h, int steps, int batch, float lr, float res_weight,
                            uint64_t seed, double* loss);
/* Mean loss of the current weights over every block of training sample k
 * (FP32 forward, res_weight of the last training), for validation. */
int cup3d_poisson_cnn_eval(cup3d_poisson_t h, int k, double* loss);

/* Introspection for tests and reports: block faces with a coarser or finer
 * neighbour (each counted from both sides). */
int64_t cup3d_poisson_num_coarse_fine_faces(cup3d_poisson_t h);

void cup3d_poisson_destroy(cup3d_poisson_t h);
/* Last error message of this handle (or of the last failed create if h is NULL). */
const char* cup3d_poisson_last_error(cup3d_poisson_t h);

#ifdef __cplusplus
}
#endif
#endif /* CUP3D_POISSON_H */
