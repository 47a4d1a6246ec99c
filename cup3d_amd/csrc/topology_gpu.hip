// Topology build on the device: the per-face tables the SpMV, the caller
// kernels and the learned warm start read (same-level neighbours, face
// kinds, coarse/fine partners, index parities, per-block h, 1/h, h^3), built
// from the block list (level, index[3]) in three kernels instead of a serial
// host pass. It replaces the face classification of the reference's
// PoissonSolverExp::getMat (source/poisson/PoissonSolverExp.cpp:221-373: the
// per-face same-level / coarser / finer tests of makeFlux, .cpp:157-219, and
// the periodic wrap of .cpp:361), which reruns after every mesh change
// (Simulation.cpp:385).
//
//   k_topo_init    clear the hash table and the info words (a fill kernel,
//                  no memset)
//   k_topo_insert  one thread per block: validate level and index, insert
//                  key(level, i, j, k) -> the smallest block id holding it
//                  (open addressing, linear probing, 64-bit CAS)
//   k_topo_blocks  one thread per block: duplicate test, h, 1/h, h^3 (FP64
//                  and FP32), index parities, the origin block
//   k_topo_faces   one thread per block face: same-level neighbour (with the
//                  periodic wrap), else the coarser block, else the four finer
//                  blocks by quadrant; a face with none of them violates the
//                  2:1 balance
// Errors are reported as the smallest offending block (or block face) id, so
// the message names the same block a sequential pass would stop at.

#include <hip/hip_runtime.h>

#include "poisson_internal.h"
#include "topology_gpu.h"

namespace cup3d {
namespace {

constexpr unsigned long long KEY_EMPTY = ~0ull;
constexpr int INDEX_BITS = 20;
constexpr int LEVEL_MAX = 15;

__device__ __forceinline__ unsigned long long tkey(int l, int i, int j, int k) {
  return ((unsigned long long)l << 60) | ((unsigned long long)i << 40) |
         ((unsigned long long)j << 20) | (unsigned long long)k;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014): spreads the packed
  // (level, i, j, k) keys over the table
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__device__ int tfind(const TopoTable& t, int l, int i, int j, int k) {
  if (l < 0 || l > LEVEL_MAX || i < 0 || j < 0 || k < 0 || i >= (1 << INDEX_BITS) ||
      j >= (1 << INDEX_BITS) || k >= (1 << INDEX_BITS))
    return -1;
  const unsigned long long key = tkey(l, i, j, k);
  unsigned long long slot = mix64(key) & t.mask;
  for (unsigned long long probe = 0; probe <= t.mask; ++probe) {
    const unsigned long long s = t.keys[slot];
    if (s == key) return t.vals[slot];
    if (s == KEY_EMPTY) return -1;
    slot = (slot + 1) & t.mask;
  }
  return -1;
}

// level 0..15 and 0 <= index < (bpd << level) <= 2^20 in every direction
__device__ __forceinline__ bool block_ok(const TopoParams& p, int4 B) {
  if (B.x < 0 || B.x > LEVEL_MAX || B.y < 0 || B.z < 0 || B.w < 0) return false;
  const int I[3] = {B.y, B.z, B.w};
  for (int d = 0; d < 3; ++d)
    if (I[d] >= (p.bpd[d] << B.x) || (p.bpd[d] << B.x) > (1 << INDEX_BITS)) return false;
  return true;
}

__global__ __launch_bounds__(WG_THREADS) void k_topo_init(TopoTable t, int* info) {
  const unsigned long long i = (unsigned long long)blockIdx.x * WG_THREADS + threadIdx.x;
  if (i <= t.mask) {
    t.keys[i] = KEY_EMPTY;
    t.vals[i] = 0x7fffffff;
  }
  if (i < TOPO_INFO_WORDS) info[i] = TopoInfoInit(i);
}

__global__ __launch_bounds__(WG_THREADS) void k_topo_insert(TopoParams p, TopoTable t,
                                                            int* info) {
  const int b = blockIdx.x * WG_THREADS + threadIdx.x;
  if (b >= p.nb) return;
  const int4 B = reinterpret_cast<const int4*>(p.blocks)[b];
  const int l = B.x, I[3] = {B.y, B.z, B.w};
  if (l < 0 || l > LEVEL_MAX || I[0] < 0 || I[1] < 0 || I[2] < 0) {
    atomicMin(&info[TI_INVALID], b);
    return;
  }
  for (int d = 0; d < 3; ++d) {
    // p.bpd[d] << l fits 32 bits (bpd < 2^16, l <= 15)
    if (I[d] >= (p.bpd[d] << l)) {
      atomicMin(&info[TI_OUTSIDE], b);
      return;
    }
    if ((p.bpd[d] << l) > (1 << INDEX_BITS)) {  // the key holds 20-bit indices
      atomicMin(&info[TI_TOO_FINE], b);
      return;
    }
  }
  const unsigned long long key = tkey(l, I[0], I[1], I[2]);
  unsigned long long slot = mix64(key) & t.mask;
  for (unsigned long long probe = 0; probe <= t.mask; ++probe) {
    const unsigned long long prev = atomicCAS(&t.keys[slot], KEY_EMPTY, key);
    if (prev == KEY_EMPTY || prev == key) {
      atomicMin(&t.vals[slot], b);
      return;
    }
    slot = (slot + 1) & t.mask;
  }
  atomicMin(&info[TI_FULL], b);  // unreachable: the table holds >= 2 nb slots
}

__global__ __launch_bounds__(WG_THREADS) void k_topo_blocks(TopoParams p, TopoTable t,
                                                            TopoOut o, int* info) {
  const int b = blockIdx.x * WG_THREADS + threadIdx.x;
  if (b >= p.nb) return;
  const int4 B = reinterpret_cast<const int4*>(p.blocks)[b];
  if (!block_ok(p, B)) return;  // reported by k_topo_insert
  if (tfind(t, B.x, B.y, B.z, B.w) != b) atomicMin(&info[TI_DUPLICATE], b);
  // h0 / 2^l is exact (a power-of-two scaling), as is its FP32 rounding of
  // the same FP64 value
  const double h = p.h0 / (double)(1 << B.x);
  const double ih = 1.0 / h;
  o.h[b] = h;
  o.invh[b] = ih;
  o.h3[b] = h * h * h;
  o.hf[b] = (float)h;
  o.invhf[b] = (float)ih;
  o.bpar[b] = (int8_t)((B.y & 1) | ((B.z & 1) << 1) | ((B.w & 1) << 2));
  o.meta[(size_t)b * 32 + 30] = o.bpar[b];
  o.meta[(size_t)b * 32 + 31] = 0;
  if (B.y == 0 && B.z == 0 && B.w == 0) atomicMax(&info[TI_ORIGIN], b);
}

__global__ __launch_bounds__(WG_THREADS) void k_topo_faces(TopoParams p, TopoTable t,
                                                           TopoOut o, int* info) {
  const int gid = blockIdx.x * WG_THREADS + threadIdx.x;
  if (gid >= p.nb * 6) return;
  const int b = gid / 6, f = gid - 6 * b;
  const int4 B = reinterpret_cast<const int4*>(p.blocks)[b];
  if (!block_ok(p, B)) return;  // reported by k_topo_insert
  const int l = B.x;
  const int d = f >> 1, s = (f & 1) ? 1 : -1;
  const int n = p.bpd[d] << l;
  int I[3] = {B.y, B.z, B.w};
  int kind = FACE_NONE, nbr = -1, part[4] = {-1, -1, -1, -1};
  const bool boundary = (s < 0) ? I[d] == 0 : I[d] == n - 1;
  if (!(boundary && p.bc[d] != 1)) {  // a wall or freespace face has no flux
    I[d] = ((I[d] + s) % n + n) % n;
    const int same = tfind(t, l, I[0], I[1], I[2]);
    if (same >= 0) {
      kind = FACE_SAME;
      nbr = same;
    } else {
      const int c = tfind(t, l - 1, I[0] >> 1, I[1] >> 1, I[2] >> 1);
      if (c >= 0) {
        kind = FACE_COARSE;
        part[0] = c;
        atomicAdd(&info[TI_N_COARSE], 1);
      } else {
        const int t1 = d == 0 ? 1 : (d == 1 ? 2 : 0), t2 = d == 0 ? 2 : (d == 1 ? 0 : 1);
        bool ok = true;
        for (int q = 0; q < 4; ++q) {  // the children adjacent to this face
          int C[3] = {2 * I[0], 2 * I[1], 2 * I[2]};
          C[d] += (s < 0) ? 1 : 0;
          C[t1] += q & 1;
          C[t2] += (q >> 1) & 1;
          part[q] = tfind(t, l + 1, C[0], C[1], C[2]);
          ok = ok && part[q] >= 0;
        }
        if (ok) {
          kind = FACE_FINE;
          atomicAdd(&info[TI_N_FINE], 1);
        } else {
          atomicMin(&info[TI_BALANCE], gid);
          for (int q = 0; q < 4; ++q) part[q] = -1;
        }
      }
    }
  }
  o.kind[gid] = (int8_t)kind;
  o.nbr[gid] = nbr;
  reinterpret_cast<int4*>(o.fnb)[gid] = make_int4(part[0], part[1], part[2], part[3]);
  reinterpret_cast<int4*>(o.meta + (size_t)b * 32)[f] = make_int4(part[0], part[1], part[2], part[3]);
  o.meta[(size_t)b * 32 + 24 + f] = kind;
}

}  // namespace

void launch_topology_build(hipStream_t s, const TopoParams& p, const TopoTable& t,
                           const TopoOut& o, int* info) {
  const unsigned long long cap = t.mask + 1;
  const int ninit = (int)((cap + WG_THREADS - 1) / WG_THREADS);
  k_topo_init<<<ninit, WG_THREADS, 0, s>>>(t, info);
  const int nwb = (p.nb + WG_THREADS - 1) / WG_THREADS;
  k_topo_insert<<<nwb, WG_THREADS, 0, s>>>(p, t, info);
  k_topo_blocks<<<nwb, WG_THREADS, 0, s>>>(p, t, o, info);
  const int nwf = (p.nb * 6 + WG_THREADS - 1) / WG_THREADS;
  k_topo_faces<<<nwf, WG_THREADS, 0, s>>>(p, t, o, info);
}

}  // namespace cup3d
