// Device topology build (topology_gpu.hip): parameters, hash table and
// output tables of the three build kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cup3d {

struct TopoParams {
  int nb;
  int bpd[3];        // base blocks per direction, each < 2^16
  int bc[3];         // 0 freespace, 1 periodic, 2 wall
  double h0;
  const int* blocks;  // [nb][4] (level, ix, iy, iz), 16-byte aligned
};

// Open-addressing table: key(level, i, j, k) -> smallest block id with it.
struct TopoTable {
  unsigned long long* keys;  // [mask + 1]
  int* vals;                 // [mask + 1]
  unsigned long long mask;   // capacity - 1, capacity a power of two >= 2 nb
};

struct TopoOut {
  int* nbr;       // [nb*6] same-level neighbour, or -1
  int8_t* kind;   // [nb*6] FaceKind
  int* fnb;       // [nb*24] coarse partner (slot 0) or finer blocks by quadrant
  int8_t* bpar;   // [nb] index parities
  double* h;      // [nb]
  double* invh;
  double* h3;
  float* hf;
  float* invhf;
  int* meta;      // [nb][32]: fnb (24), kind (6), bpar, 0 (the SpMV's one load per lane)
};

// Info words written by the build. The error words hold the smallest
// offending block id (TI_BALANCE: block * 6 + face), INT_MAX when none.
enum TopoInfo : int {
  TI_INVALID = 0,    // level outside 0..15 or a negative index
  TI_OUTSIDE = 1,    // index outside the domain at its level
  TI_TOO_FINE = 2,   // bpd << level above 2^20 (the key's index width)
  TI_DUPLICATE = 3,  // a block repeating an earlier block's level and index
  TI_BALANCE = 4,    // a face with no same-level, coarser or four finer blocks
  TI_FULL = 5,       // hash table full (cannot happen at capacity >= 2 nb)
  TI_ORIGIN = 6,     // largest block id at index (0, 0, 0), or -1
  TI_N_COARSE = 7,   // faces with a coarser neighbour
  TI_N_FINE = 8,     // faces with finer neighbours
  TOPO_INFO_WORDS = 9
};

__host__ __device__ inline int TopoInfoInit(unsigned long long i) {
  return i == TI_ORIGIN ? -1 : (i >= TI_N_COARSE ? 0 : 0x7fffffff);
}

void launch_topology_build(hipStream_t s, const TopoParams& p, const TopoTable& t,
                           const TopoOut& o, int* info);

}  // namespace cup3d
