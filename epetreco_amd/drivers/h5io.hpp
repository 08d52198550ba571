// h5io.hpp -- the reference's HDF5 file formats, through the HDF5 C API.
//
// Measurement: dataset "messung", float, all channels, any shape
// (H5Reader.hpp); the measurement list holds the nonzero channels in
// ascending channel order with their values (H5File2DefaultMeasurementList.h:37-52).
// Density: dataset "density", float, dims (nx, ny, nz) (z fastest, the voxel
// id order), scalar float attributes xmin ymin zmin xmax ymax zmax and int
// attributes xnbin ynbin znbin (H5DensityWriter.hpp:106-202; H5DensityReader.hpp).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "epet.h"
#include "hdf5.h"

namespace epet_h5 {

inline void h5ok(long long rc, const std::string& what) {
  if (rc < 0) throw std::runtime_error("HDF5: " + what);
}

inline std::vector<float> read_floats(const std::string& fn, const char* dataset) {
  const hid_t f = H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  h5ok(f, "cannot open " + fn);
  const hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
  if (ds < 0) {
    H5Fclose(f);
    throw std::runtime_error("HDF5: no dataset " + std::string(dataset) + " in " + fn);
  }
  const hid_t sp = H5Dget_space(ds);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<float> v(n > 0 ? (size_t)n : 0);
  herr_t rc = 0;
  if (n > 0) rc = H5Dread(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  H5Sclose(sp);
  H5Dclose(ds);
  H5Fclose(f);
  h5ok(rc, "read " + std::string(dataset) + " from " + fn);
  return v;
}

inline std::vector<float> read_measurement(const std::string& fn) {
  return read_floats(fn, "messung");
}

inline std::vector<float> read_density(const std::string& fn) {
  return read_floats(fn, "density");
}

inline void write_dataset(const std::string& fn, const char* name, const float* x, int nd,
                          const hsize_t* dims, const epet_grid* g) {
  const hid_t f = H5Fcreate(fn.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  h5ok(f, "cannot create " + fn);
  const hid_t sp = H5Screate_simple(nd, dims, nullptr);
  const hid_t ds = H5Dcreate2(f, name, H5T_NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t rc = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, x);
  if (rc >= 0 && g) {
    const char* lo[3] = {"xmin", "ymin", "zmin"};
    const char* hi[3] = {"xmax", "ymax", "zmax"};
    const char* nb[3] = {"xnbin", "ynbin", "znbin"};
    const hid_t sc = H5Screate(H5S_SCALAR);
    for (int d = 0; d < 3 && rc >= 0; d++) {
      const float mn = g->o[d], mx = g->o[d] + g->n[d] * g->d[d];
      const int nbin = g->n[d];
      const hid_t a0 = H5Acreate2(ds, lo[d], H5T_NATIVE_FLOAT, sc, H5P_DEFAULT, H5P_DEFAULT);
      rc = H5Awrite(a0, H5T_NATIVE_FLOAT, &mn);
      H5Aclose(a0);
      const hid_t a1 = H5Acreate2(ds, hi[d], H5T_NATIVE_FLOAT, sc, H5P_DEFAULT, H5P_DEFAULT);
      if (rc >= 0) rc = H5Awrite(a1, H5T_NATIVE_FLOAT, &mx);
      H5Aclose(a1);
      const hid_t a2 = H5Acreate2(ds, nb[d], H5T_NATIVE_INT, sc, H5P_DEFAULT, H5P_DEFAULT);
      if (rc >= 0) rc = H5Awrite(a2, H5T_NATIVE_INT, &nbin);
      H5Aclose(a2);
    }
    H5Sclose(sc);
  }
  if (ds >= 0) H5Dclose(ds);
  H5Sclose(sp);
  H5Fclose(f);
  h5ok(rc, "write " + std::string(name) + " to " + fn);
}

inline void write_measurement(const std::string& fn, const std::vector<float>& v) {
  const hsize_t dims[1] = {(hsize_t)v.size()};
  write_dataset(fn, "messung", v.data(), 1, dims, nullptr);
}

inline void write_density(const std::string& fn, const float* x, const epet_grid& g) {
  const hsize_t dims[3] = {(hsize_t)g.n[0], (hsize_t)g.n[1], (hsize_t)g.n[2]};
  write_dataset(fn, "density", x, 3, dims, &g);
}

// one scalar attribute of dataset "density"
template <typename T>
inline T density_attribute(const std::string& fn, const char* name) {
  const hid_t f = H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  h5ok(f, "cannot open " + fn);
  const hid_t ds = H5Dopen2(f, "density", H5P_DEFAULT);
  const hid_t a = ds < 0 ? -1 : H5Aopen(ds, name, H5P_DEFAULT);
  T v{};
  herr_t rc = a < 0 ? -1 : H5Aread(a, sizeof(T) == 4 && T(0.5) != T(0) ? H5T_NATIVE_FLOAT
                                                                      : H5T_NATIVE_INT, &v);
  if (a >= 0) H5Aclose(a);
  if (ds >= 0) H5Dclose(ds);
  H5Fclose(f);
  h5ok(rc, std::string("attribute ") + name);
  return v;
}

// nonzero channels in ascending order and their values
inline void measurement_list(const std::vector<float>& raw, std::vector<int>& cnl,
                             std::vector<float>& val) {
  cnl.clear();
  val.clear();
  for (size_t i = 0; i < raw.size(); i++)
    if (raw[i] != 0.0f) {
      cnl.push_back((int)i);
      val.push_back(raw[i]);
    }
}

}  // namespace epet_h5
