/*
 * cup3d_h5.h — block fields on disk in CubismUP_3D's multi-block HDF5 layout
 * (libcup3d_h5.so, host only).
 *
 * The reference dumps its fields with Cubism's DumpHDF5_MPI2
 * (source/Simulation.cpp:245-258) and reads them back in its conversion tool
 * (tools/uniform_conversion/main.cpp:59-99 get_amr_groups, :102-150
 * get_amr_dataset): <base>.h5 holds dataset "dset", every group's cells
 * concatenated (x fastest, then y, then z); <base>-groups.h5 holds "origins"
 * (ox, oy, oz, h per group) and "indices" (nx, ny, nz, index[3], level per
 * group). These functions write the solver's fields in that layout and read
 * such dumps into the solver's block layout ([nblocks][8][8][8] doubles,
 * [iz][iy][ix]; block list nblocks x 4 ints: level, index[3]), so a dumped
 * pressure or rhs can be loaded into cup3d_poisson_* and results written
 * back for the reference's tools.
 *
 * Errors: < 0 with a NUL-terminated message in err (errlen bytes).
 */
#ifndef CUP3D_H5_H
#define CUP3D_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* group_x >= 1: blocks per group at most (runs of blocks that follow each
 * other in the list and along x at one level); float32 != 0 stores FP32
 * (what the reference's reader reads), else FP64. */
int cup3d_h5_write_blocks(const char* base, int64_t nblocks, const int32_t* blocks, double h0,
                          const double* values, int float32, int group_x, char* err,
                          int errlen);
/* Blocks held by the groups of <base>-groups.h5. */
int64_t cup3d_h5_count_blocks(const char* base, char* err, int errlen);
/* Read <base>: the blocks of each group in x-fastest order; *h0 = the level-0
 * cell size the groups imply. Returns the number of blocks. */
int64_t cup3d_h5_read_blocks(const char* base, int64_t cap, int32_t* blocks, double* values,
                             double* h0, char* err, int errlen);

#ifdef __cplusplus
}
#endif
#endif /* CUP3D_H5_H */
