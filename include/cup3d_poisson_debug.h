/*
 * cup3d_poisson_debug.h — host-only diagnostics of libcup3d_poisson.
 *
 * No counterpart in the reference. The reference assembles its operator as an
 * explicit COO matrix (PoissonSolverExp::getMat, PoissonSolverExp.cpp:221-373);
 * this library applies it matrix-free. cup3d_debug_operator_csr() writes the
 * operator exactly as the SpMV kernel applies it (7-point stencil over
 * same-level faces plus the merged coarse-fine interface rows, bMean = 3 row
 * included; the bMean = 1 row replacement and the bMean = 2 rank-one term
 * are applied at run time and are not part of it) as CSR with sorted columns, built on the host from the same
 * topology tables the device uses. It needs no GPU.
 *
 * Call with cap = 0 to get nnz; then with arrays of n+1 / nnz / nnz entries.
 * Returns nnz, or < 0 on invalid input.
 */
#ifndef CUP3D_POISSON_DEBUG_H
#define CUP3D_POISSON_DEBUG_H

#include <stdint.h>

#include "cup3d_poisson.h"

#ifdef __cplusplus
extern "C" {
#endif

int64_t cup3d_debug_operator_csr(const cup3d_poisson_config* cfg, int64_t nblocks,
                                 const cup3d_block_desc* blocks, int64_t* rowptr, int32_t* col,
                                 double* val, int64_t cap);

/* y = A x with the FP32 operator of the mixed-precision inner solve (x cast
 * to FP32, y cast back; FP64 device vectors), for checking it against the
 * FP64 operator. */
int cup3d_debug_apply_operator32(cup3d_poisson_t h, const double* d_x, double* d_y);

/* Kernel timing for roofline reports: runs one kernel class `reps` times on
 * the handle's stream between two HIP events, on the data of the last solve,
 * and returns the average duration in ms. which: 0 = operator (SpMV),
 * 1 = block-Jacobi GEMM, 2 = fused x/r update (K5), 3 = fused p update +
 * GEMM (K1), 4 = SpMV + dot (K2), in the handle's precision. Also returns
 * the algorithmic bytes (vectors read + written) and flops of one launch
 * (GEMM flops count the MFMA passes actually issued).
 * Must follow a solve on the current topology. Returns 0 or < 0. */
int cup3d_debug_time_kernel(cup3d_poisson_t h, int which, int reps, double* avg_ms,
                            double* bytes, double* flops);

/* The same for the implicit-diffusion kernels (cup3d_diffusion_*) of velocity
 * component 0 with 1 / (nu dt): which 0 = Helmholtz SpMV, 1 = DST block
 * inverse fused with the p update (the K1 form), 2 = DST block inverse alone;
 * algorithmic bytes 2, 5 and 2 FP64 vectors per launch. Needs a topology;
 * overwrites the handle's FP64 Krylov vectors. */
int cup3d_debug_time_diffusion_kernel(cup3d_poisson_t h, int which, double nu, double dt,
                                      int reps, double* avg_ms, double* bytes, double* flops);

/* In-solve kernel timing of the mixed-precision form. While enabled, every
 * solve on the handle records HIP events on its stream around the main kernel
 * of K1 (p update + GEMM), K2 (SpMV + (rhat, nu)), K3 (s update + GEMM), K4
 * (SpMV + (t, s), (t, t)) and K5 (x/r update), their one-workgroup finalize
 * kernels excluded, in the middle iteration of each polled batch of
 * iterations (one batch = 8 iterations by default), counted when the batch
 * ran past it.
 * cup3d_debug_profile(h, 1) enables and clears, (h, 0) disables;
 * cup3d_debug_profile_read returns the average durations in microseconds
 * (avg_us[5]: K1 .. K5) and the number of samples, and clears them.
 * The FP64 form returns < 0. */
int cup3d_debug_profile(cup3d_poisson_t h, int enable);
/* Device-to-device streaming copy rate in bytes/s (read + written) of a
 * nontemporal float4 copy between two buffers of `bytes` on `device`,
 * averaged over `reps` copies: the practical HBM ceiling reported next to
 * the datasheet peak. Returns < 0 on error. */
double cup3d_debug_copy_rate(int device, int64_t bytes, int reps);
int cup3d_debug_profile_read(cup3d_poisson_t h, double* avg_us, long long* samples);

/* Full-batch FP32 gradient (n_weights floats) and loss of the learned warm
 * start over every block of training sample k (grids up to 4096 blocks), for
 * checking the backward kernels against an FP64 reference. */
int cup3d_debug_cnn_gradient(cup3d_poisson_t h, int k, float res_weight, float* grad,
                             double* loss);

/* Topology tables of a block list, as the host build of this library makes
 * them (the build the explicit-matrix export above uses): same-level
 * neighbours nbr[nb*6] (-1: none), face kinds kind[nb*6] (0 none, 1 same
 * level, 2 coarser, 3 finer), partners fnb[nb*24] (coarser block in slot 0,
 * or the four finer blocks by quadrant), index parities bpar[nb], h, 1/h and
 * h^3 per block; info[3] = origin block, coarser faces, finer faces. Any
 * output may be NULL. On a rejected block list returns < 0 and copies the
 * message into err (errlen bytes, NUL-terminated). Needs no GPU. */
int cup3d_debug_host_topology(const cup3d_poisson_config* cfg, int64_t nblocks,
                              const cup3d_block_desc* blocks, int32_t* nbr, int8_t* kind,
                              int32_t* fnb, int8_t* bpar, double* h, double* invh, double* h3,
                              int64_t* info, char* err, int errlen);

/* The same tables as the handle's device topology build made them for the
 * current topology (copied to host arrays of the sizes above; any may be
 * NULL), for comparing the two builds. Returns 0 or < 0. */
int cup3d_debug_device_topology(cup3d_poisson_t h, int32_t* nbr, int8_t* kind, int32_t* fnb,
                                int8_t* bpar, double* hh, double* invh, double* h3,
                                int64_t* info);

#ifdef __cplusplus
}
#endif
#endif
