// Learned warm start: network, inference and training kernels.
//
// Network (per 8^3 block, one-cell halo), in the increment form of the
// pressure solve (PressureProjection.cpp:368-385 solves for p - p_old with
// rhs b - A p_old):
//   input  2 x 10^3      ch0 = res / (h_b s_r), ch1 = p_prev / s_p,
//                        res = b - A p_prev, s_r = RMS(res / h_b), s_p = RMS(p_prev)
//   conv 3^3 valid  2 -> 16, GELU
//   conv 3^3 same  16 -> 16, GELU      (zero padding inside the block)
//   conv 3^3 same  16 -> 1
//   p_net = p_prev + s_r * out
// res / h_b has the units of the 7-point stencil applied to the increment
// (A = h^3 lap), so the input is commensurate on every AMR level. Halo: a face
// halo cell takes the adjacent cell of the same-level neighbour block; every
// other halo cell (wall, coarse/fine face, edge, corner) repeats the nearest
// cell of the block.
//
// Inference (k_cnn_infer): one workgroup per block, activations in LDS as
// FP16 channel-last [8][8][8][16] (the zero padding selected on read), the
// input as the [10][10][10] patch, each layer an implicit
// GEMM on v_mfma_f32_16x16x32_f16 (M = 16 cells, N = 16 output channels,
// K = 27 offsets x input channels; the 8 consecutive K of a lane are 8
// channels of one offset: one 16-byte LDS read), FP32 accumulation, bias and
// GELU in FP32.
// Training (k_cnn_train_item + k_cnn_adam): FP32, one workgroup per
// (sample, block) item, forward and backward with activations in global
// scratch, per-item gradients summed in item order (deterministic), Adam.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>

#include "cnn.h"
#include "device_common.h"
#include "kernels.h"

namespace cup3d {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
}

// Index into a field of the cell that padded position (pz, py, px) in 0..9
// of block b takes its value from (the halo rule above).
__device__ __forceinline__ size_t halo_index(const int* __restrict__ nbr, int b, int pz, int py,
                                             int px) {
  int c[3] = {px - 1, py - 1, pz - 1};
  int out = 0, dim = -1;
  for (int d = 0; d < 3; ++d)
    if (c[d] < 0 || c[d] > 7) {
      ++out;
      dim = d;
    }
  int blk = b;
  if (out == 1) {
    const int face = 2 * dim + (c[dim] > 7 ? 1 : 0);
    const int nb = nbr[b * 6 + face];
    if (nb >= 0) {
      blk = nb;
      c[dim] = c[dim] > 7 ? 0 : 7;
    }
  }
  for (int d = 0; d < 3; ++d) c[d] = c[d] < 0 ? 0 : (c[d] > 7 ? 7 : c[d]);
  return (size_t)blk * BLEN + c[2] * 64 + c[1] * 8 + c[0];
}

// Value of field f at padded position (pz, py, px) in 0..9 of block b.
template <typename T>
__device__ __forceinline__ float halo_value(const T* __restrict__ f, const int* __restrict__ nbr,
                                            int b, int pz, int py, int px) {
  return (float)f[halo_index(nbr, b, pz, py, px)];
}

// ---------------------------------------------------------------------------
// scal[slot] = RMS(f / h_b) (h = nullptr: RMS(f)); deterministic two-pass sum
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG_THREADS) void k_cnn_rms(long long n, const double* __restrict__ f,
                                                        const double* __restrict__ h,
                                                        double* part) {
  double v[1] = {0.0};
  const long long base = (long long)blockIdx.x * (4 * BLEN);
  for (int e = threadIdx.x; e < 4 * BLEN; e += WG_THREADS) {
    const long long i = base + e;
    if (i < n) {
      const double x = h ? f[i] / h[i / BLEN] : f[i];
      v[0] += x * x;
    }
  }
  grid_store_partial<1>(v, part);
}

__global__ __launch_bounds__(WG_THREADS) void k_cnn_rms_fin(long long n, double* scal, int slot,
                                                            const double* part, int nparts) {
  double tot[1];
  grid_sum_partials<1>(part, nparts, tot);
  if (threadIdx.x == 0) {
    const double s = sqrt(tot[0] / (double)n);
    scal[slot] = s > 0 ? s : 1.0;
  }
}

// scal[2] = volume-weighted mean of p - p_prev: sum h^3 (p - p_prev) / sum h^3
__global__ __launch_bounds__(WG_THREADS) void k_cnn_dmean(long long n, const double* __restrict__ p,
                                                          const double* __restrict__ prev,
                                                          const double* __restrict__ h,
                                                          double* part) {
  double v[2] = {0.0, 0.0};
  const long long base = (long long)blockIdx.x * (4 * BLEN);
  for (int e = threadIdx.x; e < 4 * BLEN; e += WG_THREADS) {
    const long long i = base + e;
    if (i < n) {
      const double hb = h[i / BLEN], h3 = hb * hb * hb;
      v[0] += h3 * (p[i] - prev[i]);
      v[1] += h3;
    }
  }
  grid_store_partial<2>(v, part);
}

__global__ __launch_bounds__(WG_THREADS) void k_cnn_dmean_fin(double* scal, const double* part,
                                                              int nparts) {
  double tot[2];
  grid_sum_partials<2>(part, nparts, tot);
  if (threadIdx.x == 0) scal[2] = tot[1] > 0 ? tot[0] / tot[1] : 0.0;
}

// ---------------------------------------------------------------------------
// FP16 weight packing for the implicit-GEMM inference.
// Fragment kt, lane l (n = l % 16 output channel, g = l / 16), element e:
//   layer 1: offset o = 4 kt + g, channel c = e (c < 2 real)
//   layers 2, 3: offset o = 2 kt + g / 2, channel c = 8 (g % 2) + e
// (o = dz*9 + dy*3 + dx; o >= 27 and missing channels are zero; layer 3 has
// a single real output channel n = 0).
// ---------------------------------------------------------------------------
__global__ void k_cnn_pack(const float* __restrict__ w, _Float16* __restrict__ pk) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= CNN_PK) return;
  const int e = idx & 7, l = (idx >> 3) & 63, frag = idx >> 9;
  const int n = l & 15, g = l >> 4;
  float v = 0.f;
  if (frag < CNN_KT1) {
    const int o = 4 * frag + g, c = e;
    if (o < 27 && c < 2) v = w[(n * 2 + c) * 27 + o];
  } else if (frag < CNN_KT1 + CNN_KT2) {
    const int kt = frag - CNN_KT1, o = 2 * kt + g / 2, c = 8 * (g % 2) + e;
    if (o < 27) v = w[CNN_OW2 + (n * CNN_C + c) * 27 + o];
  } else {
    const int kt = frag - CNN_KT1 - CNN_KT2, o = 2 * kt + g / 2, c = 8 * (g % 2) + e;
    if (o < 27 && n == 0) v = w[CNN_OW3 + c * 27 + o];
  }
  pk[idx] = (_Float16)v;
}

// ---------------------------------------------------------------------------
// Inference: one workgroup (4 waves) per block.
// ---------------------------------------------------------------------------
constexpr int PADC = 1000;  // 10^3 padded cells

__device__ __forceinline__ int padded(int m) {  // cell 0..511 -> padded index
  return (((m >> 6) + 1) * 10 + ((m >> 3) & 7) + 1) * 10 + (m & 7) + 1;
}

// One conv layer as implicit GEMM, K-steps KT; output tile handler f(m, n,
// value) for each of this wave's 8 M-tiles. L1: `in` is the 10^3 input patch
// (8 halfs per cell, valid conv); otherwise `in` holds the block's 512 cells
// (CNN_C halfs each, no halo) and reads outside the block are the zero
// padding of the "same" conv (selected, not stored).
template <int KT, bool L1, typename F>
__device__ __forceinline__ void conv_layer(const _Float16* __restrict__ in,
                                           const half8_t* __restrict__ wfrag, int w, int l,
                                           F&& f) {
  const int lr = l & 15, g = l >> 4;
  half8_t bq[KT];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt) bq[kt] = wfrag[kt * 64 + l];
  half8_t zero;
#pragma unroll
  for (int e = 0; e < 8; ++e) zero[e] = (_Float16)0.f;
  for (int t = 0; t < 8; ++t) {
    const int tile = 8 * w + t;
    const int m = 16 * tile + lr;  // this lane's A row (cell)
    const int mz = m >> 6, my = (m >> 3) & 7, mx = m & 7;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      const int o = L1 ? 4 * kt + g : 2 * kt + (g >> 1);
      const int c0 = L1 ? 0 : 8 * (g & 1);
      half8_t a = zero;
      if (o < 27) {
        const int dz = o / 9, dy = (o / 3) % 3, dx = o % 3;
        if (L1) {
          const int p = ((mz + dz) * 10 + (my + dy)) * 10 + (mx + dx);
          a = *reinterpret_cast<const half8_t*>(&in[p * 8]);
        } else {
          const int z = mz + dz - 1, y = my + dy - 1, x = mx + dx - 1;
          const bool inside = (unsigned)z < 8u && (unsigned)y < 8u && (unsigned)x < 8u;
          const int p = inside ? (z * 8 + y) * 8 + x : 0;
          const half8_t v = *reinterpret_cast<const half8_t*>(&in[p * CNN_C + c0]);
          a = inside ? v : zero;
        }
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bq[kt], acc, 0, 0, 0);
    }
    // register q of lane l holds D[4*(l/16) + q][l % 16]
#pragma unroll
    for (int q = 0; q < 4; ++q) f(16 * tile + 4 * g + q, lr, acc[q]);
  }
}

// LDS of the inference: the input patch (1000 cells x 8 halfs), later the
// second layer's activations (512 x 16), and the first layer's (512 x 16):
// 32 KiB per workgroup, five workgroups per CU.
constexpr int CNN_IN_HALFS = PADC * 8 > BLEN * CNN_C ? PADC * 8 : BLEN * CNN_C;

__global__ __launch_bounds__(WG_THREADS) void k_cnn_infer(int nb, const int* __restrict__ nbr,
                                                          const double* __restrict__ h,
                                                          const double* __restrict__ res,
                                                          const double* __restrict__ prev,
                                                          const double* __restrict__ scal,
                                                          const _Float16* __restrict__ wpk,
                                                          const float* __restrict__ w,
                                                          double* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) _Float16 bufIn[CNN_IN_HALFS];  // input, then act2
  __shared__ __attribute__((aligned(16))) _Float16 bufA[BLEN * CNN_C];   // act1
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  if (b >= nb) return;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
  const double s_r = scal[0];
  const float inv_sp = (float)(1.0 / scal[1]);
  const float inv_hs = (float)(1.0 / (h[b] * s_r));
  // input: 8-channel cells (2 real); the indices of a thread's (up to 4)
  // positions first, then all 8 loads, so the gather costs two memory round
  // trips, not one per position
  constexpr int NP = (PADC + WG_THREADS - 1) / WG_THREADS;
  size_t gi[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = tid + j * WG_THREADS;
    gi[j] = p < PADC ? halo_index(nbr, b, p / 100, (p / 10) % 10, p % 10) : (size_t)b * BLEN;
  }
  double rv[NP], pv[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    rv[j] = res[gi[j]];
    pv[j] = prev[gi[j]];
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = tid + j * WG_THREADS;
    if (p < PADC) {
      half8_t v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = (_Float16)0.f;
      v[0] = (_Float16)((float)rv[j] * inv_hs);
      v[1] = (_Float16)((float)pv[j] * inv_sp);
      *reinterpret_cast<half8_t*>(&bufIn[p * 8]) = v;
    }
  }
  __syncthreads();
  const half8_t* wf = reinterpret_cast<const half8_t*>(wpk);
  conv_layer<CNN_KT1, true>(bufIn, wf, wv, l, [&](int m, int n, float v) {
    bufA[m * CNN_C + n] = (_Float16)gelu(v + w[CNN_OB1 + n]);
  });
  __syncthreads();
  conv_layer<CNN_KT2, false>(bufA, wf + CNN_KT1 * 64, wv, l, [&](int m, int n, float v) {
    bufIn[m * CNN_C + n] = (_Float16)gelu(v + w[CNN_OB2 + n]);
  });
  __syncthreads();
  const float b3 = w[CNN_OB3];
  const size_t off = (size_t)b * BLEN;
  conv_layer<CNN_KT3, false>(bufIn, wf + (CNN_KT1 + CNN_KT2) * 64, wv, l,
                             [&](int m, int n, float v) {
                               if (n == 0) out[off + m] = prev[off + m] + s_r * (double)(v + b3);
                             });
}

// ---------------------------------------------------------------------------
// Training (FP32)
// ---------------------------------------------------------------------------
struct SampleRef {
  const float* res_h;
  const float* prev;
  const float* target;
  float s_r;
  float s_p;
};

// per-item scratch layout (floats)
constexpr int TS_X = 0;                        // [2][1000]
constexpr int TS_Z1 = TS_X + 2 * PADC;         // [16][512]
constexpr int TS_H1 = TS_Z1 + CNN_C * BLEN;    // [16][1000] padded
constexpr int TS_Z2 = TS_H1 + CNN_C * PADC;    // [16][512]
constexpr int TS_H2 = TS_Z2 + CNN_C * BLEN;    // [16][1000] padded
constexpr int TS_D = TS_H2 + CNN_C * PADC;     // [512] dout
constexpr int TS_DZ2 = TS_D + BLEN;            // [16][512]
constexpr int TS_DZ1 = TS_DZ2 + CNN_C * BLEN;  // [16][512]
constexpr int TS_SIZE = TS_DZ1 + CNN_C * BLEN;

// Forward (and, unless EVAL, backward) of one (sample, block) item.
template <bool EVAL>
__global__ __launch_bounds__(WG_THREADS) void k_cnn_train_item(
    int nitems, const int* __restrict__ items, const SampleRef* __restrict__ samples,
    const int* __restrict__ nbr, const float* __restrict__ wg, float* __restrict__ scratch,
    float* __restrict__ gbuf, float* __restrict__ lossbuf, float inv_ntot, float lam) {
  __shared__ float W[CNN_NW];
  __shared__ float OP[PADC];   // target halo, then the prediction inside it
  __shared__ float RHO[BLEN];  // local residual
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const int tid = threadIdx.x;
  const int k = items[2 * it], b = items[2 * it + 1];
  const SampleRef S = samples[k];
  float* __restrict__ sc = scratch + (size_t)it * TS_SIZE;
  float* X = sc + TS_X;
  float* Z1 = sc + TS_Z1;
  float* H1 = sc + TS_H1;
  float* Z2 = sc + TS_Z2;
  float* H2 = sc + TS_H2;
  float* D = sc + TS_D;
  float* DZ2 = sc + TS_DZ2;
  float* DZ1 = sc + TS_DZ1;
  for (int j = tid; j < CNN_NW; j += WG_THREADS) W[j] = wg[j];
  const float inv_sr = 1.0f / S.s_r, inv_sp = 1.0f / S.s_p;
  for (int p = tid; p < PADC; p += WG_THREADS) {
    const int pz = p / 100, py = (p / 10) % 10, px = p % 10;
    X[p] = halo_value(S.res_h, nbr, b, pz, py, px) * inv_sr;
    X[PADC + p] = halo_value(S.prev, nbr, b, pz, py, px) * inv_sp;
    OP[p] = halo_value(S.target, nbr, b, pz, py, px);
    const bool ring = pz == 0 || pz == 9 || py == 0 || py == 9 || px == 0 || px == 9;
    if (ring)
      for (int c = 0; c < CNN_C; ++c) {
        H1[c * PADC + p] = 0.f;
        H2[c * PADC + p] = 0.f;
      }
  }
  __syncthreads();
  // layer 1 (valid)
  for (int m = tid; m < BLEN; m += WG_THREADS) {
    const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
    for (int o = 0; o < CNN_C; ++o) {
      float a = W[CNN_OB1 + o];
      for (int c = 0; c < 2; ++c)
        for (int off = 0; off < 27; ++off) {
          const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
          a += W[(o * 2 + c) * 27 + off] * X[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx];
        }
      Z1[o * BLEN + m] = a;
      H1[o * PADC + padded(m)] = gelu(a);
    }
  }
  __syncthreads();
  // layer 2 (same)
  for (int m = tid; m < BLEN; m += WG_THREADS) {
    const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
    for (int o = 0; o < CNN_C; ++o) {
      float a = W[CNN_OB2 + o];
      for (int c = 0; c < CNN_C; ++c)
        for (int off = 0; off < 27; ++off) {
          const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
          a += W[CNN_OW2 + (o * CNN_C + c) * 27 + off] *
               H1[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx];
        }
      Z2[o * BLEN + m] = a;
      H2[o * PADC + padded(m)] = gelu(a);
    }
  }
  __syncthreads();
  // layer 3 (same) and the loss gradient
  __shared__ float lred[WG_THREADS];
  float lsum = 0.f;
  const float* T = S.target + (size_t)b * BLEN;
  for (int m = tid; m < BLEN; m += WG_THREADS) {
    const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
    float a = W[CNN_OB3];
    for (int c = 0; c < CNN_C; ++c)
      for (int off = 0; off < 27; ++off) {
        const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
        a += W[CNN_OW3 + c * 27 + off] * H2[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx];
      }
    const float d = a - T[m];
    lsum += d * d;
    D[m] = d * inv_ntot;
    OP[padded(m)] = a;
  }
  __syncthreads();
  if (lam != 0.f) {
    // rho = lap7(out) - q, the stencil closed by the target halo outside
    // the block; its gradient adds lam * (in-block L^T rho) / N to dL/dout
    for (int m = tid; m < BLEN; m += WG_THREADS) {
      const int P = padded(m);
      const float r = ((OP[P - 1] + OP[P + 1]) + (OP[P - 10] + OP[P + 10]) +
                       (OP[P - 100] + OP[P + 100])) - 6.f * OP[P] - X[P];
      RHO[m] = r;
      lsum += lam * r * r;
    }
    __syncthreads();
    for (int m = tid; m < BLEN; m += WG_THREADS) {
      const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
      float gsum = -6.f * RHO[m];
      if (x > 0) gsum += RHO[m - 1];
      if (x < 7) gsum += RHO[m + 1];
      if (y > 0) gsum += RHO[m - 8];
      if (y < 7) gsum += RHO[m + 8];
      if (z > 0) gsum += RHO[m - 64];
      if (z < 7) gsum += RHO[m + 64];
      D[m] += lam * gsum * inv_ntot;
    }
  }
  lred[tid] = lsum;
  __syncthreads();  // lred and D complete before the sum and the backward pass
  if (tid == 0) {
    float s = 0.f;
    for (int i = 0; i < WG_THREADS; ++i) s += lred[i];
    lossbuf[it] = s;
  }
  if (EVAL) return;
  float* G = gbuf + (size_t)it * CNN_NW;
  // dW3, db3
  for (int j = tid; j < CNN_NW3; j += WG_THREADS) {
    const int c = j / 27, off = j % 27, dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) {
      const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
      a += H2[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx] * D[m];
    }
    G[CNN_OW3 + j] = a;
  }
  if (tid == 0) {
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) a += D[m];
    G[CNN_OB3] = a;
  }
  // dZ2 = (W3^T * D) . gelu'(Z2)
  for (int i = tid; i < CNN_C * BLEN; i += WG_THREADS) {
    const int c = i / BLEN, m = i % BLEN, z = m >> 6, y = (m >> 3) & 7, x = m & 7;
    float a = 0.f;
    for (int off = 0; off < 27; ++off) {
      const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
      const int zz = z + 1 - dz, yy = y + 1 - dy, xx = x + 1 - dx;
      if (zz >= 0 && zz < 8 && yy >= 0 && yy < 8 && xx >= 0 && xx < 8)
        a += W[CNN_OW3 + c * 27 + off] * D[(zz * 8 + yy) * 8 + xx];
    }
    DZ2[i] = a * gelu_grad(Z2[i]);
  }
  __syncthreads();
  // dW2, db2
  for (int j = tid; j < CNN_NW2; j += WG_THREADS) {
    const int o = j / (CNN_C * 27), c = (j / 27) % CNN_C, off = j % 27;
    const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) {
      const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
      a += H1[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx] * DZ2[o * BLEN + m];
    }
    G[CNN_OW2 + j] = a;
  }
  for (int o = tid; o < CNN_C; o += WG_THREADS) {
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) a += DZ2[o * BLEN + m];
    G[CNN_OB2 + o] = a;
  }
  // dZ1 = (W2^T * dZ2) . gelu'(Z1)
  for (int i = tid; i < CNN_C * BLEN; i += WG_THREADS) {
    const int c = i / BLEN, m = i % BLEN, z = m >> 6, y = (m >> 3) & 7, x = m & 7;
    float a = 0.f;
    for (int o = 0; o < CNN_C; ++o)
      for (int off = 0; off < 27; ++off) {
        const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
        const int zz = z + 1 - dz, yy = y + 1 - dy, xx = x + 1 - dx;
        if (zz >= 0 && zz < 8 && yy >= 0 && yy < 8 && xx >= 0 && xx < 8)
          a += W[CNN_OW2 + (o * CNN_C + c) * 27 + off] * DZ2[o * BLEN + (zz * 8 + yy) * 8 + xx];
      }
    DZ1[i] = a * gelu_grad(Z1[i]);
  }
  __syncthreads();
  // dW1, db1
  for (int j = tid; j < CNN_NW1; j += WG_THREADS) {
    const int o = j / 54, c = (j / 27) % 2, off = j % 27;
    const int dz = off / 9, dy = (off / 3) % 3, dx = off % 3;
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) {
      const int z = m >> 6, y = (m >> 3) & 7, x = m & 7;
      a += X[c * PADC + ((z + dz) * 10 + y + dy) * 10 + x + dx] * DZ1[o * BLEN + m];
    }
    G[j] = a;
  }
  for (int o = tid; o < CNN_C; o += WG_THREADS) {
    float a = 0.f;
    for (int m = 0; m < BLEN; ++m) a += DZ1[o * BLEN + m];
    G[CNN_OB1 + o] = a;
  }
}

// Sum the per-item gradients in item order and take one Adam step
// (Kingma & Ba 2015; beta1 = 0.9, beta2 = 0.999, eps = 1e-8).
__global__ void k_cnn_adam(int nitems, const float* __restrict__ gbuf, float* __restrict__ w,
                           float* __restrict__ mv, float lr, float bc1, float bc2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= CNN_NW) return;
  float g = 0.f;
  for (int i = 0; i < nitems; ++i) g += gbuf[(size_t)i * CNN_NW + j];
  float m = mv[j], v = mv[CNN_NW + j];
  m = 0.9f * m + 0.1f * g;
  v = 0.999f * v + 0.001f * g * g;
  mv[j] = m;
  mv[CNN_NW + j] = v;
  w[j] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// Training-set entry: res / h_b, p_prev and (p - p_prev - m) / s_r in FP32,
// m the volume-weighted mean of p - p_prev (scal[2]): a constant shift of the
// pressure is invisible to the operator and to the inputs, so it is not a
// target.
__global__ void k_cnn_sample(long long n, const double* __restrict__ res,
                             const double* __restrict__ prev, const double* __restrict__ p,
                             const double* __restrict__ h, const double* __restrict__ scal,
                             float* __restrict__ rh, float* __restrict__ pv,
                             float* __restrict__ tg) {
  const double inv_s = 1.0 / scal[0], m = scal[2];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    rh[i] = (float)(res[i] / h[i / BLEN]);
    pv[i] = (float)prev[i];
    tg[i] = (float)((p[i] - prev[i] - m) * inv_s);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define CCHECK(expr)                                            \
  do {                                                          \
    hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) {                                     \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);  \
      return -1;                                                \
    }                                                           \
  } while (0)

int cnn_num_weights() { return CNN_NW; }

int cnn_init(CnnState& C, std::string& err) {
  CCHECK(hipMalloc(&C.d_w, CNN_NW * sizeof(float)));
  CCHECK(hipMalloc(&C.d_adam, 2 * CNN_NW * sizeof(float)));
  CCHECK(hipMalloc(&C.d_wpk, CNN_PK * sizeof(_Float16)));
  CCHECK(hipMalloc(&C.d_scal, 8 * sizeof(double)));
  CCHECK(hipMemset(C.d_w, 0, CNN_NW * sizeof(float)));
  CCHECK(hipMemset(C.d_adam, 0, 2 * CNN_NW * sizeof(float)));
  return 0;
}

int cnn_set_topology(CnnState& C, int nb, const int* d_nbr, const double* d_h, std::string& err) {
  cnn_clear_samples(C);
  C.nb = nb;
  C.n = (long long)nb * BLEN;
  C.d_nbr = d_nbr;
  C.d_h = d_h;
  if (nb <= C.cap_nb && C.d_part && C.d_pred) return 0;  // the buffers grow only
  if (C.d_part) (void)hipFree(C.d_part);
  if (C.d_pred) (void)hipFree(C.d_pred);
  C.d_part = nullptr;
  C.d_pred = nullptr;
  C.cap_nb = 0;
  // partial sums: up to 2 per workgroup of 4 blocks
  CCHECK(hipMalloc(&C.d_part, (2 * ((nb + 3) / 4) + 16) * sizeof(double)));
  CCHECK(hipMalloc(&C.d_pred, C.n * sizeof(double)));
  C.cap_nb = nb;
  return 0;
}

static int repack(CnnState& C, hipStream_t s, std::string& err) {
  k_cnn_pack<<<(CNN_PK + 255) / 256, 256, 0, s>>>(C.d_w, static_cast<_Float16*>(C.d_wpk));
  CCHECK(hipGetLastError());
  return 0;
}

int cnn_set_weights(CnnState& C, const float* w, int nw, std::string& err) {
  if (nw != CNN_NW || !w) {
    err = "cnn_set_weights: expected " + std::to_string(CNN_NW) + " weights";
    return -1;
  }
  CCHECK(hipMemcpy(C.d_w, w, CNN_NW * sizeof(float), hipMemcpyHostToDevice));
  CCHECK(hipMemset(C.d_adam, 0, 2 * CNN_NW * sizeof(float)));
  C.adam_t = 0;
  if (repack(C, nullptr, err)) return -1;
  CCHECK(hipDeviceSynchronize());
  C.have_weights = true;
  return 0;
}

int cnn_get_weights(CnnState& C, float* w, int nw, std::string& err) {
  if (nw != CNN_NW || !w) {
    err = "cnn_get_weights: expected " + std::to_string(CNN_NW) + " weights";
    return -1;
  }
  CCHECK(hipDeviceSynchronize());
  CCHECK(hipMemcpy(w, C.d_w, CNN_NW * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// scal[0] = s_r = RMS(res / h_b), scal[1] = s_p = RMS(p_prev)
static int scales(CnnState& C, hipStream_t s, const double* res, const double* prev,
                  std::string& err) {
  const int nwg = (int)((C.n + 4 * BLEN - 1) / (4 * BLEN));
  k_cnn_rms<<<nwg, WG_THREADS, 0, s>>>(C.n, res, C.d_h, C.d_part);
  k_cnn_rms_fin<<<1, WG_THREADS, 0, s>>>(C.n, C.d_scal, 0, C.d_part, nwg);
  k_cnn_rms<<<nwg, WG_THREADS, 0, s>>>(C.n, prev, nullptr, C.d_part);
  k_cnn_rms_fin<<<1, WG_THREADS, 0, s>>>(C.n, C.d_scal, 1, C.d_part, nwg);
  CCHECK(hipGetLastError());
  return 0;
}

int cnn_predict(CnnState& C, hipStream_t s, const double* d_res, const double* d_prev,
                double* d_out, std::string& err) {
  if (!C.have_weights) {
    err = "cnn_predict: no weights (set or train them first)";
    return -1;
  }
  if (scales(C, s, d_res, d_prev, err)) return -1;
  k_cnn_infer<<<C.nb, WG_THREADS, 0, s>>>(C.nb, C.d_nbr, C.d_h, d_res, d_prev, C.d_scal,
                                          static_cast<const _Float16*>(C.d_wpk), C.d_w, d_out);
  CCHECK(hipGetLastError());
  CCHECK(hipStreamSynchronize(s));
  return 0;
}

int cnn_add_sample(CnnState& C, hipStream_t s, const double* d_res, const double* d_prev,
                   const double* d_p, std::string& err) {
  if (C.nb <= 0) {
    err = "cnn_add_sample: no topology";
    return -1;
  }
  CnnSample S;
  CCHECK(hipMalloc(&S.res_h, 3 * C.n * sizeof(float)));
  S.prev = S.res_h + C.n;
  S.target = S.prev + C.n;
  if (scales(C, s, d_res, d_prev, err)) return -1;
  const int nwg = (int)((C.n + 4 * BLEN - 1) / (4 * BLEN));
  k_cnn_dmean<<<nwg, WG_THREADS, 0, s>>>(C.n, d_p, d_prev, C.d_h, C.d_part);
  k_cnn_dmean_fin<<<1, WG_THREADS, 0, s>>>(C.d_scal, C.d_part, nwg);
  k_cnn_sample<<<4096, 256, 0, s>>>(C.n, d_res, d_prev, d_p, C.d_h, C.d_scal, S.res_h, S.prev,
                                    S.target);
  CCHECK(hipGetLastError());
  double sc[2] = {1.0, 1.0};
  CCHECK(hipMemcpyAsync(sc, C.d_scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  CCHECK(hipStreamSynchronize(s));
  S.s_r = (float)sc[0];
  S.s_p = (float)sc[1];
  C.samples.push_back(S);
  return 0;
}

void cnn_clear_samples(CnnState& C) {
  for (auto& S : C.samples) (void)hipFree(S.res_h);
  C.samples.clear();
}

static int ensure_train_buffers(CnnState& C, int nitems, std::string& err) {
  const long long need = (long long)nitems * (TS_SIZE + CNN_NW + 1);
  if (need > C.scratch_cap) {
    if (C.d_scratch) (void)hipFree(C.d_scratch);
    C.d_scratch = nullptr;
    CCHECK(hipMalloc(&C.d_scratch, need * sizeof(float)));
    C.scratch_cap = need;
  }
  if (nitems > C.items_cap) {
    if (C.d_items) (void)hipFree(C.d_items);
    C.d_items = nullptr;
    CCHECK(hipMalloc(&C.d_items, 2 * nitems * sizeof(int)));
    C.items_cap = nitems;
  }
  return 0;
}

static int upload_samples(CnnState& C, std::vector<SampleRef>& refs, SampleRef** d_refs,
                          std::string& err) {
  refs.resize(C.samples.size());
  for (size_t k = 0; k < C.samples.size(); ++k)
    refs[k] = {C.samples[k].res_h, C.samples[k].prev, C.samples[k].target, C.samples[k].s_r,
               C.samples[k].s_p};
  CCHECK(hipMalloc(d_refs, refs.size() * sizeof(SampleRef)));
  CCHECK(hipMemcpy(*d_refs, refs.data(), refs.size() * sizeof(SampleRef), hipMemcpyHostToDevice));
  return 0;
}

int cnn_train(CnnState& C, hipStream_t s, int steps, int batch, float lr, float res_weight,
              uint64_t seed, double* loss, std::string& err) {
  if (C.samples.empty()) {
    err = "cnn_train: empty training set";
    return -1;
  }
  if (steps <= 0 || batch <= 0 || batch > 65536 || !(res_weight >= 0.f)) {
    err = "cnn_train: bad steps/batch/res_weight";
    return -1;
  }
  C.res_weight = res_weight;
  if (!C.have_weights) {  // He-style initialisation from the seed
    std::mt19937_64 R(seed ^ 0x5eedULL);
    std::vector<float> w(CNN_NW, 0.f);
    auto fill = [&](int off, int n, int fan_in) {
      const double sd = std::sqrt(2.0 / fan_in);
      for (int i = 0; i < n; ++i) {
        const double u1 = (double((R() >> 11) + 1)) * (1.0 / 9007199254740993.0);
        const double u2 = double(R() >> 11) * (1.0 / 9007199254740992.0);
        w[off + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2));
      }
    };
    fill(0, CNN_NW1, 2 * 27);
    fill(CNN_OW2, CNN_NW2, CNN_C * 27);
    fill(CNN_OW3, CNN_NW3, CNN_C * 27);
    for (int i = CNN_OW3; i < CNN_OW3 + CNN_NW3; ++i) w[i] *= 0.1f;  // start near p_prev
    if (cnn_set_weights(C, w.data(), CNN_NW, err)) return -1;
  }
  if (ensure_train_buffers(C, batch, err)) return -1;
  std::vector<SampleRef> refs;
  SampleRef* d_refs = nullptr;
  if (upload_samples(C, refs, &d_refs, err)) return -1;
  float* scratch = C.d_scratch;
  float* gbuf = scratch + (size_t)batch * TS_SIZE;
  float* lbuf = gbuf + (size_t)batch * CNN_NW;
  std::vector<int> items(2 * batch);
  std::vector<float> lh(batch);
  std::mt19937_64 R(seed);
  const int ns = (int)C.samples.size();
  const float inv_ntot = 1.0f / (float)((double)batch * BLEN);
  for (int step = 0; step < steps; ++step) {
    for (int i = 0; i < batch; ++i) {
      items[2 * i] = (int)(R() % (uint64_t)ns);
      items[2 * i + 1] = (int)(R() % (uint64_t)C.nb);
    }
    CCHECK(hipMemcpyAsync(C.d_items, items.data(), items.size() * sizeof(int),
                          hipMemcpyHostToDevice, s));
    k_cnn_train_item<false><<<batch, WG_THREADS, 0, s>>>(batch, C.d_items, d_refs, C.d_nbr, C.d_w,
                                                         scratch, gbuf, lbuf, inv_ntot,
                                                         C.res_weight);
    C.adam_t += 1;
    const float bc1 = 1.0f - std::pow(0.9f, (float)C.adam_t);
    const float bc2 = 1.0f - std::pow(0.999f, (float)C.adam_t);
    k_cnn_adam<<<(CNN_NW + 255) / 256, 256, 0, s>>>(batch, gbuf, C.d_w, C.d_adam, lr, bc1, bc2);
    if (step == 0 || step == steps - 1) {
      CCHECK(hipMemcpyAsync(lh.data(), lbuf, batch * sizeof(float), hipMemcpyDeviceToHost, s));
      CCHECK(hipStreamSynchronize(s));
      double L = 0;
      for (float v : lh) L += v;
      L *= 0.5 / ((double)batch * BLEN);
      if (loss) loss[step == 0 ? 0 : 1] = L;
      if (loss && steps == 1) loss[1] = L;
    }
  }
  CCHECK(hipGetLastError());
  if (repack(C, s, err)) return -1;
  CCHECK(hipStreamSynchronize(s));
  (void)hipFree(d_refs);
  return 0;
}

int cnn_eval(CnnState& C, hipStream_t s, int k, double* loss, std::string& err) {
  if (k < 0 || k >= (int)C.samples.size()) {
    err = "cnn_eval: no such sample";
    return -1;
  }
  const int chunk = std::min(C.nb, 2048);
  if (ensure_train_buffers(C, chunk, err)) return -1;
  std::vector<SampleRef> refs;
  SampleRef* d_refs = nullptr;
  if (upload_samples(C, refs, &d_refs, err)) return -1;
  float* scratch = C.d_scratch;
  float* gbuf = scratch + (size_t)chunk * TS_SIZE;
  float* lbuf = gbuf + (size_t)chunk * CNN_NW;
  std::vector<int> items(2 * chunk);
  std::vector<float> lh(chunk);
  double L = 0;
  for (int b0 = 0; b0 < C.nb; b0 += chunk) {
    const int m = std::min(chunk, C.nb - b0);
    for (int i = 0; i < m; ++i) {
      items[2 * i] = k;
      items[2 * i + 1] = b0 + i;
    }
    CCHECK(hipMemcpyAsync(C.d_items, items.data(), 2 * m * sizeof(int), hipMemcpyHostToDevice, s));
    k_cnn_train_item<true><<<m, WG_THREADS, 0, s>>>(m, C.d_items, d_refs, C.d_nbr, C.d_w,
                                                    scratch, gbuf, lbuf, 0.f, C.res_weight);
    CCHECK(hipMemcpyAsync(lh.data(), lbuf, m * sizeof(float), hipMemcpyDeviceToHost, s));
    CCHECK(hipStreamSynchronize(s));
    for (int i = 0; i < m; ++i) L += lh[i];
  }
  (void)hipFree(d_refs);
  *loss = 0.5 * L / ((double)C.nb * BLEN);
  return 0;
}

__global__ void k_cnn_sum_grads(int nitems, const float* __restrict__ gbuf, float* __restrict__ g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= CNN_NW) return;
  float a = 0.f;
  for (int i = 0; i < nitems; ++i) a += gbuf[(size_t)i * CNN_NW + j];
  g[j] = a;
}

int cnn_gradient(CnnState& C, hipStream_t s, int k, float res_weight, float* grad, double* loss,
                 std::string& err) {
  if (k < 0 || k >= (int)C.samples.size()) {
    err = "cnn_gradient: no such sample";
    return -1;
  }
  if (!C.have_weights) {
    err = "cnn_gradient: no weights";
    return -1;
  }
  if (C.nb > 4096) {
    err = "cnn_gradient: grid too large for a full-batch gradient";
    return -1;
  }
  const int m = C.nb;
  if (ensure_train_buffers(C, m + 1, err)) return -1;
  std::vector<SampleRef> refs;
  SampleRef* d_refs = nullptr;
  if (upload_samples(C, refs, &d_refs, err)) return -1;
  float* scratch = C.d_scratch;
  float* gbuf = scratch + (size_t)(m + 1) * TS_SIZE;
  float* lbuf = gbuf + (size_t)(m + 1) * CNN_NW;
  float* gsum = gbuf + (size_t)m * CNN_NW;
  std::vector<int> items(2 * m);
  for (int i = 0; i < m; ++i) {
    items[2 * i] = k;
    items[2 * i + 1] = i;
  }
  CCHECK(hipMemcpyAsync(C.d_items, items.data(), 2 * m * sizeof(int), hipMemcpyHostToDevice, s));
  k_cnn_train_item<false><<<m, WG_THREADS, 0, s>>>(m, C.d_items, d_refs, C.d_nbr, C.d_w, scratch,
                                                   gbuf, lbuf, 1.0f / ((float)m * BLEN),
                                                   res_weight);
  k_cnn_sum_grads<<<(CNN_NW + 255) / 256, 256, 0, s>>>(m, gbuf, gsum);
  std::vector<float> lh(m);
  CCHECK(hipMemcpyAsync(grad, gsum, CNN_NW * sizeof(float), hipMemcpyDeviceToHost, s));
  CCHECK(hipMemcpyAsync(lh.data(), lbuf, m * sizeof(float), hipMemcpyDeviceToHost, s));
  CCHECK(hipStreamSynchronize(s));
  double L = 0;
  for (float v : lh) L += v;
  *loss = 0.5 * L / ((double)m * BLEN);
  (void)hipFree(d_refs);
  return 0;
}

void cnn_free(CnnState& C) {
  cnn_clear_samples(C);
  (void)hipFree(C.d_w);
  (void)hipFree(C.d_adam);
  (void)hipFree(C.d_wpk);
  (void)hipFree(C.d_scal);
  if (C.d_part) (void)hipFree(C.d_part);
  if (C.d_pred) (void)hipFree(C.d_pred);
  if (C.d_scratch) (void)hipFree(C.d_scratch);
  if (C.d_items) (void)hipFree(C.d_items);
  C = CnnState();
}

}  // namespace cup3d
