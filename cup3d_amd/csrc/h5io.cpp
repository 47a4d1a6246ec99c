// Block fields on disk in CubismUP_3D's multi-block HDF5 layout, so that
// fields a simulation dumped (the pressure, an rhs) can be loaded into the
// solver's block layout and solver results written back for the reference's
// tools. The layout is the one the reference's own reader consumes
// (tools/uniform_conversion/main.cpp:59-99 get_amr_groups, :102-150
// get_amr_dataset), written by Cubism's DumpHDF5_MPI2 at every dump
// (Simulation.cpp:245-258; Cubism itself is absent from the reference mount):
//   <base>.h5         dataset "dset": every group's cells, concatenated, each
//                     group x fastest, then y, then z
//                     (value = dset[start_g + x + y nx + z nx ny])
//   <base>-groups.h5  dataset "origins": 4 doubles per group (ox, oy, oz, h)
//                     dataset "indices": 7 ints per group
//                                        (nx, ny, nz, index[3], level)
// A group is a box of nx x ny x nz cells (multiples of 8) at one level whose
// lowest block has block index index[3]; the writer here emits runs of up to
// `group_x` blocks that follow each other in the block list and along x.
// Values are stored as FP32 (what the reference's reader reads) or FP64.
//
// Host code only; serial HDF5 (the image's HDF5 1.10 has no MPI-IO).

#include <hdf5.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int BS = 8;

struct Group {
  double o[3], h;
  int n[3], index[3], level;
};

void set_err(char* err, int errlen, const std::string& m) {
  if (err && errlen > 0) {
    std::strncpy(err, m.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = '\0';
  }
}

// RAII for HDF5 ids (H5Fclose / H5Dclose / H5Sclose).
struct Hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Hid() {
    if (id >= 0 && close) close(id);
  }
  Hid(const Hid&) = delete;
  Hid& operator=(const Hid&) = delete;
};

bool write_1d(hid_t file, const char* name, hid_t file_type, hid_t mem_type, const void* data,
              hsize_t n) {
  Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose);
  if (space.id < 0) return false;
  Hid ds(H5Dcreate2(file, name, file_type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
         H5Dclose);
  if (ds.id < 0) return false;
  return n == 0 || H5Dwrite(ds.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

bool read_1d(hid_t file, const char* name, hid_t mem_type, size_t elem, std::vector<char>& out,
             hsize_t& n) {
  Hid ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (ds.id < 0) return false;
  Hid space(H5Dget_space(ds.id), H5Sclose);
  if (space.id < 0 || H5Sget_simple_extent_ndims(space.id) != 1) return false;
  if (H5Sget_simple_extent_dims(space.id, &n, nullptr) < 0) return false;
  out.assign((size_t)n * elem, 0);
  return n == 0 || H5Dread(ds.id, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

bool read_groups(const std::string& base, std::vector<Group>& G, std::string& msg) {
  Hid f(H5Fopen((base + "-groups.h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) {
    msg = "cannot open " + base + "-groups.h5";
    return false;
  }
  std::vector<char> o, ix;
  hsize_t no = 0, ni = 0;
  if (!read_1d(f.id, "origins", H5T_NATIVE_DOUBLE, sizeof(double), o, no) ||
      !read_1d(f.id, "indices", H5T_NATIVE_INT, sizeof(int), ix, ni)) {
    msg = "datasets origins / indices missing or not 1-D";
    return false;
  }
  if (no % 4 || ni % 7 || no / 4 != ni / 7) {
    msg = "origins / indices sizes do not describe the same groups";
    return false;
  }
  const double* od = reinterpret_cast<const double*>(o.data());
  const int* id = reinterpret_cast<const int*>(ix.data());
  G.resize(no / 4);
  for (size_t g = 0; g < G.size(); ++g) {
    for (int d = 0; d < 3; ++d) {
      G[g].o[d] = od[4 * g + d];
      G[g].n[d] = id[7 * g + d];
      G[g].index[d] = id[7 * g + 3 + d];
    }
    G[g].h = od[4 * g + 3];
    G[g].level = id[7 * g + 6];
    for (int d = 0; d < 3; ++d)
      if (G[g].n[d] <= 0 || G[g].n[d] % BS || G[g].index[d] < 0) {
        msg = "group " + std::to_string(g) + ": sizes must be positive multiples of 8";
        return false;
      }
    if (G[g].level < 0 || G[g].level > 30) {
      msg = "group " + std::to_string(g) + ": invalid level";
      return false;
    }
  }
  return true;
}

}  // namespace

extern "C" {

// Write a scalar field given in the solver's block layout ([nblocks][8][8][8],
// [iz][iy][ix]) with its block list (nblocks x 4 ints: level, index[3]) and
// level-0 cell size h0. group_x >= 1: blocks per group at most (runs of
// blocks consecutive in the list and along x at one level). float32 != 0
// stores FP32. Returns 0, or < 0 with a message in err.
int cup3d_h5_write_blocks(const char* base, int64_t nblocks, const int32_t* blocks, double h0,
                          const double* values, int float32, int group_x, char* err,
                          int errlen) {
  if (!base || nblocks <= 0 || !blocks || !values || !(h0 > 0) || group_x < 1) {
    set_err(err, errlen, "invalid arguments");
    return -1;
  }
  // groups: runs of blocks following each other in the list and along x
  std::vector<Group> G;
  std::vector<int64_t> first;  // first block of each group
  for (int64_t b = 0; b < nblocks; ++b) {
    const int32_t* B = blocks + 4 * b;
    if (B[0] < 0 || B[1] < 0 || B[2] < 0 || B[3] < 0) {
      set_err(err, errlen, "block " + std::to_string(b) + ": invalid level/index");
      return -1;
    }
    if (!G.empty()) {
      Group& g = G.back();
      const int32_t* P = blocks + 4 * (b - 1);
      if (g.n[0] / BS < group_x && B[0] == g.level && P[0] == B[0] && B[1] == P[1] + 1 &&
          B[2] == P[2] && B[3] == P[3]) {
        g.n[0] += BS;
        continue;
      }
    }
    Group g;
    g.level = B[0];
    g.h = h0 / double(1 << B[0]);
    for (int d = 0; d < 3; ++d) {
      g.index[d] = B[1 + d];
      g.n[d] = BS;
      g.o[d] = B[1 + d] * BS * g.h;
    }
    G.push_back(g);
    first.push_back(b);
  }
  std::vector<double> origins(4 * G.size());
  std::vector<int> indices(7 * G.size());
  std::vector<double> dset((size_t)nblocks * BS * BS * BS);
  size_t pos = 0;
  for (size_t k = 0; k < G.size(); ++k) {
    const Group& g = G[k];
    for (int d = 0; d < 3; ++d) {
      origins[4 * k + d] = g.o[d];
      indices[7 * k + d] = g.n[d];
      indices[7 * k + 3 + d] = g.index[d];
    }
    origins[4 * k + 3] = g.h;
    indices[7 * k + 6] = g.level;
    for (int z = 0; z < BS; ++z)
      for (int y = 0; y < BS; ++y)
        for (int x = 0; x < g.n[0]; ++x) {
          const int64_t b = first[k] + x / BS;
          dset[pos++] = values[(size_t)b * 512 + z * 64 + y * 8 + (x % BS)];
        }
  }
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are reported through err
  {
    Hid f(H5Fcreate((std::string(base) + ".h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
          H5Fclose);
    if (f.id < 0 ||
        !write_1d(f.id, "dset", float32 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE,
                  dset.data(), dset.size())) {
      set_err(err, errlen, std::string("cannot write ") + base + ".h5");
      return -1;
    }
  }
  {
    Hid f(H5Fcreate((std::string(base) + "-groups.h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT,
                    H5P_DEFAULT),
          H5Fclose);
    if (f.id < 0 ||
        !write_1d(f.id, "origins", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, origins.data(),
                  origins.size()) ||
        !write_1d(f.id, "indices", H5T_STD_I32LE, H5T_NATIVE_INT, indices.data(),
                  indices.size())) {
      set_err(err, errlen, std::string("cannot write ") + base + "-groups.h5");
      return -1;
    }
  }
  return 0;
}

// Number of 8^3 blocks the groups of <base>-groups.h5 hold, or < 0.
int64_t cup3d_h5_count_blocks(const char* base, char* err, int errlen) {
  if (!base) return -1;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  std::vector<Group> G;
  std::string msg;
  if (!read_groups(base, G, msg)) {
    set_err(err, errlen, msg);
    return -1;
  }
  int64_t nb = 0;
  for (const Group& g : G) nb += (int64_t)(g.n[0] / BS) * (g.n[1] / BS) * (g.n[2] / BS);
  return nb;
}

// Read a dump into the solver's layout: blocks (cap x 4 ints: level,
// index[3], the blocks of each group in x-fastest order), values (cap x 512
// doubles, [iz][iy][ix]) and the level-0 cell size implied by the groups
// (h * 2^level, which must agree between groups). Returns the number of
// blocks, or < 0.
int64_t cup3d_h5_read_blocks(const char* base, int64_t cap, int32_t* blocks, double* values,
                             double* h0, char* err, int errlen) {
  if (!base || !blocks || !values) return -1;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  std::vector<Group> G;
  std::string msg;
  if (!read_groups(base, G, msg)) {
    set_err(err, errlen, msg);
    return -1;
  }
  int64_t nb = 0;
  size_t ncell = 0;
  for (const Group& g : G) {
    nb += (int64_t)(g.n[0] / BS) * (g.n[1] / BS) * (g.n[2] / BS);
    ncell += (size_t)g.n[0] * g.n[1] * g.n[2];
  }
  if (nb > cap) {
    set_err(err, errlen, "more blocks than cap (" + std::to_string(nb) + ")");
    return -1;
  }
  Hid f(H5Fopen((std::string(base) + ".h5").c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) {
    set_err(err, errlen, std::string("cannot open ") + base + ".h5");
    return -1;
  }
  std::vector<char> raw;
  hsize_t n = 0;
  if (!read_1d(f.id, "dset", H5T_NATIVE_DOUBLE, sizeof(double), raw, n) || n != ncell) {
    set_err(err, errlen, "dataset dset missing or of the wrong size");
    return -1;
  }
  const double* d = reinterpret_cast<const double*>(raw.data());
  double h0v = -1.0;
  int64_t b = 0;
  size_t start = 0;
  for (size_t k = 0; k < G.size(); ++k) {
    const Group& g = G[k];
    const double h0g = g.h * double(1 << g.level);
    if (h0v < 0) h0v = h0g;
    if (std::abs(h0g - h0v) > 1e-12 * h0v) {
      set_err(err, errlen, "group " + std::to_string(k) + ": h does not match its level");
      return -1;
    }
    const int nbx = g.n[0] / BS, nby = g.n[1] / BS, nbz = g.n[2] / BS;
    for (int bz = 0; bz < nbz; ++bz)
      for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx, ++b) {
          int32_t* B = blocks + 4 * b;
          B[0] = g.level;
          B[1] = g.index[0] + bx;
          B[2] = g.index[1] + by;
          B[3] = g.index[2] + bz;
          double* out = values + (size_t)b * 512;
          for (int z = 0; z < BS; ++z)
            for (int y = 0; y < BS; ++y)
              for (int x = 0; x < BS; ++x) {
                const size_t X = (size_t)bx * BS + x, Y = (size_t)by * BS + y,
                             Z = (size_t)bz * BS + z;
                out[z * 64 + y * 8 + x] = d[start + X + Y * g.n[0] + Z * g.n[0] * g.n[1]];
              }
        }
    start += (size_t)g.n[0] * g.n[1] * g.n[2];
  }
  if (h0) *h0 = h0v;
  return nb;
}

}  // extern "C"
