// CPU ORACLE — TEST INFRASTRUCTURE ONLY (the checker of libcup3d_h5; only
// tests/ load it).
//
// A serial restatement of the reference's reader of CubismUP_3D's
// multi-block HDF5 dumps, tools/uniform_conversion/main.cpp:
//   get_amr_groups  (:59-99)   origins / indices of <base>-groups.h5 into
//                              groups (origin, h, nx, ny, nz, index, level);
//   get_amr_dataset (:102-150) dataset "dset" of <base>.h5 read as FP32 in
//                              8 hyperslab chunks;
//   convert_to_uniform (:152-238, the arithmetic; one rank, no MPI) the
//                              uniform grid at 1/Cfactor of the finest
//                              resolution: each cell's value times
//                              1/Cfactor^3, replicated over the finest cells
//                              it covers and summed into the uniform cell
//                              holding each of them.
// The product's reader and writer (cup3d_amd/csrc/h5io.cpp) are checked
// against it: what the product writes, the reference's own reader reads back
// as the same blocks and values.
//
// Pinned by: a constant field converts to the same constant; on a uniform
// level-0 grid every uniform cell is its block's mean (tests/test_h5io.py).

#include <hdf5.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

namespace {

constexpr int BS = 8;
constexpr int CFACTOR = 8;

struct BlockGroup {  // main.cpp:20-27
  double h;
  int nx, ny, nz;
  double ox, oy, oz;
  int level;
  int index[3];
};

std::vector<BlockGroup> get_amr_groups(const std::string& filename) {
  hid_t file_id = H5Fopen((filename + "-groups.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file_id < 0) return {};
  hid_t dataset_origins = H5Dopen2(file_id, "origins", H5P_DEFAULT);
  hid_t dataset_indices = H5Dopen2(file_id, "indices", H5P_DEFAULT);
  hsize_t dim_origins, dim_indices;
  hid_t fspace_origins = H5Dget_space(dataset_origins);
  hid_t fspace_indices = H5Dget_space(dataset_indices);
  H5Sget_simple_extent_dims(fspace_origins, &dim_origins, NULL);
  H5Sget_simple_extent_dims(fspace_indices, &dim_indices, NULL);
  std::vector<double> origins(dim_origins);
  std::vector<int> indices(dim_indices);
  H5Dread(dataset_origins, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, origins.data());
  H5Dread(dataset_indices, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, indices.data());
  H5Dclose(dataset_origins);
  H5Dclose(dataset_indices);
  H5Sclose(fspace_origins);
  H5Sclose(fspace_indices);
  H5Fclose(file_id);
  std::vector<BlockGroup> groups(dim_origins / 4);
  for (size_t i = 0; i < groups.size(); i++) {
    groups[i].ox = origins[4 * i];
    groups[i].oy = origins[4 * i + 1];
    groups[i].oz = origins[4 * i + 2];
    groups[i].h = origins[4 * i + 3];
    groups[i].nx = indices[7 * i];
    groups[i].ny = indices[7 * i + 1];
    groups[i].nz = indices[7 * i + 2];
    groups[i].index[0] = indices[7 * i + 3];
    groups[i].index[1] = indices[7 * i + 4];
    groups[i].index[2] = indices[7 * i + 5];
    groups[i].level = indices[7 * i + 6];
  }
  return groups;
}

std::vector<float> get_amr_dataset(const std::string& filename) {
  hid_t file_id = H5Fopen((filename + ".h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file_id < 0) return {};
  hid_t dataset_id = H5Dopen2(file_id, "dset", H5P_DEFAULT);
  hsize_t dim;
  hid_t fspace_id = H5Dget_space(dataset_id);
  H5Sget_simple_extent_dims(fspace_id, &dim, NULL);
  std::vector<float> amr(dim);
  const int max_chunk = 8;  // :120-140
  for (int chunk = 0; chunk < max_chunk; chunk++) {
    hsize_t count[1] = {dim / max_chunk};
    hsize_t base_tmp[1] = {chunk * count[0]};
    hid_t mspace_id = H5Screate_simple(1, count, NULL);
    H5Sselect_hyperslab(fspace_id, H5S_SELECT_SET, base_tmp, NULL, count, NULL);
    H5Dread(dataset_id, H5T_NATIVE_FLOAT, mspace_id, fspace_id, H5P_DEFAULT,
            amr.data() + base_tmp[0]);
    H5Sclose(mspace_id);
  }
  H5Dclose(dataset_id);
  H5Sclose(fspace_id);
  H5Fclose(file_id);
  return amr;
}

}  // namespace

extern "C" {

// Groups of <base>-groups.h5: origins (4 doubles) and indices (7 ints) per
// group, as get_amr_groups stores them. Returns the group count (or -1).
int64_t orc_h5_groups(const char* base, double* origins, int32_t* indices, int64_t cap) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::vector<BlockGroup> G = get_amr_groups(base);
  if ((int64_t)G.size() > cap) return -1;
  for (size_t i = 0; i < G.size(); ++i) {
    const BlockGroup& g = G[i];
    const double o[4] = {g.ox, g.oy, g.oz, g.h};
    const int32_t ix[7] = {g.nx, g.ny, g.nz, g.index[0], g.index[1], g.index[2], g.level};
    std::copy(o, o + 4, origins + 4 * i);
    std::copy(ix, ix + 7, indices + 7 * i);
  }
  return (int64_t)G.size();
}

// dset as FP32, as get_amr_dataset reads it. Returns its length (or -1).
int64_t orc_h5_dataset(const char* base, float* out, int64_t cap) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::vector<float> a = get_amr_dataset(base);
  if ((int64_t)a.size() > cap) return -1;
  std::copy(a.begin(), a.end(), out);
  return (int64_t)a.size();
}

// convert_to_uniform (main.cpp:152-238) on one rank: dims[3] receives the
// uniform grid's size (x, y, z); out (cap floats, [z][y][x]) its values.
int64_t orc_h5_to_uniform(const char* base, int32_t* dims, float* out, int64_t cap) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::vector<float> amr = get_amr_dataset(base);
  const std::vector<BlockGroup> allGroups = get_amr_groups(base);
  if (allGroups.empty()) return -1;
  std::vector<size_t> gbase(allGroups.size());
  gbase[0] = 0;
  for (size_t i = 1; i < allGroups.size(); i++)
    gbase[i] = gbase[i - 1] +
               (size_t)allGroups[i - 1].nx * allGroups[i - 1].ny * allGroups[i - 1].nz;
  int levelMax = -1;
  int points[3] = {0, 0, 0};
  for (const BlockGroup& g : allGroups) levelMax = std::max(g.level, levelMax);
  levelMax++;
  for (const BlockGroup& g : allGroups) {
    const int aux = 1 << (levelMax - 1 - g.level);
    points[0] = std::max(points[0], (g.index[0] * BS + g.nx) * aux);
    points[1] = std::max(points[1], (g.index[1] * BS + g.ny) * aux);
    points[2] = std::max(points[2], (g.index[2] * BS + g.nz) * aux);
  }
  const int my_start = 0, my_end = points[2];
  const size_t unc =
      (size_t)(my_end - my_start) / CFACTOR * (points[1] / CFACTOR) * (points[0] / CFACTOR);
  if ((int64_t)unc > cap) return -1;
  std::vector<float> uniform_grid(unc, 0.0f);
  const double coef = 1.0 / CFACTOR / CFACTOR / CFACTOR;
  for (size_t i = 0; i < allGroups.size(); i++) {
    const BlockGroup& group = allGroups[i];
    const int aux = 1 << (levelMax - 1 - group.level);
    const int start_x = group.index[0] * BS * aux;
    const int start_y = group.index[1] * BS * aux;
    const int start_z = group.index[2] * BS * aux;
    for (int z = 0; z < group.nz; z++)
      for (int y = 0; y < group.ny; y++)
        for (int x = 0; x < group.nx; x++) {
          const float value =
              coef * amr[gbase[i] + x + y * group.nx + (size_t)z * group.nx * group.ny];
          for (int z_up = aux * z; z_up < aux * (z + 1); z_up++)
            for (int y_up = aux * y; y_up < aux * (y + 1); y_up++)
              for (int x_up = aux * x; x_up < aux * (x + 1); x_up++) {
                const int uniform_z = start_z + z_up;
                const int uniform_y = start_y + y_up;
                const int uniform_x = start_x + x_up;
                const size_t b = uniform_x / CFACTOR +
                                 (size_t)(uniform_y / CFACTOR) * (points[0] / CFACTOR) +
                                 (size_t)((uniform_z - my_start) / CFACTOR) *
                                     (points[0] / CFACTOR) * (points[1] / CFACTOR);
                uniform_grid[b] += value;
              }
        }
  }
  for (int d = 0; d < 3; ++d) dims[d] = points[d] / CFACTOR;
  std::copy(uniform_grid.begin(), uniform_grid.end(), out);
  return (int64_t)unc;
}

}  // extern "C"
