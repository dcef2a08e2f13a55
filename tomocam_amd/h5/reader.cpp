// HDF5 projection I/O for the real-data front end (host code).
// Reference behaviour: tomocam::h5::Reader / Writer (src/hdf5/reader.h,
// src/hdf5/writer.h): float32 datasets, projections stored [nproj][nslices]
// [ncols], output written as a rank-3 float32 dataset.
#include <hdf5.h>

#include <stdexcept>
#include <string>

#include "tomocam_h5.h"

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
  int code;
  H5Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void require(bool ok, const std::string& msg) {
  if (!ok) throw H5Error(1, msg);
}

// closes an HDF5 handle with the matching function on scope exit
struct Handle {
  hid_t id;
  herr_t (*close)(hid_t);
  Handle(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) {
    if (id < 0) throw H5Error(2, what);
  }
  ~Handle() {
    if (id >= 0) close(id);
  }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
};

template <class F>
int guarded(F&& f) {
  // errors are reported through the status and tc_h5_last_error(), not
  // HDF5's automatic stderr dump
  static const bool quiet = [] {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    return true;
  }();
  (void)quiet;
  try {
    f();
    return 0;
  } catch (const H5Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
}

struct Dataset {
  Handle file, dset, space;
  int ndim = 0;
  hsize_t dims[3] = {0, 0, 0};
  Dataset(const char* fname, const char* name)
      : file(H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose,
             std::string("cannot open HDF5 file ") + fname),
        dset(H5Dopen2(file.id, name, H5P_DEFAULT), H5Dclose,
             std::string("cannot open dataset ") + name),
        space(H5Dget_space(dset.id), H5Sclose, "cannot get dataspace") {
    ndim = H5Sget_simple_extent_ndims(space.id);
    require(ndim >= 1 && ndim <= 3,
            std::string("dataset ") + name + " must have rank 1, 2 or 3");
    H5Sget_simple_extent_dims(space.id, dims, nullptr);
  }
  void require_float(const char* name) {
    Handle t(H5Dget_type(dset.id), H5Tclose, "cannot get data type");
    require(H5Tequal(t.id, H5T_NATIVE_FLOAT) > 0,
            std::string("data type mismatch: dataset ") + name +
                " is not float32");
  }
};

// rank-3 hyperslab [begin, end) along `axis` (0 or 1), read contiguously
void read_block(const char* fname, const char* name, long begin, long end,
                int axis, float* out) {
  require(fname && name && out, "null argument");
  Dataset d(fname, name);
  require(d.ndim == 3, std::string("dataset ") + name + " is not 3D");
  d.require_float(name);
  const long n = (long)d.dims[axis];
  if (end < 0) end = n;
  require(begin >= 0 && begin <= end && end <= n,
          "slice range [" + std::to_string(begin) + ", " +
              std::to_string(end) + ") out of bounds (" + std::to_string(n) +
              " slices)");
  hsize_t start[3] = {0, 0, 0}, count[3] = {d.dims[0], d.dims[1], d.dims[2]};
  start[axis] = (hsize_t)begin;
  count[axis] = (hsize_t)(end - begin);
  if (count[axis] == 0) return;
  require(H5Sselect_hyperslab(d.space.id, H5S_SELECT_SET, start, nullptr,
                              count, nullptr) >= 0,
          "hyperslab selection failed");
  Handle mem(H5Screate_simple(3, count, nullptr), H5Sclose,
             "cannot create memory space");
  if (H5Dread(d.dset.id, H5T_NATIVE_FLOAT, mem.id, d.space.id, H5P_DEFAULT,
              out) < 0)
    throw H5Error(2, std::string("reading ") + name + " failed");
}

}  // namespace

extern "C" {

int tc_h5_dims(const char* file, const char* dataset, int* ndim, long* dims) {
  return guarded([&] {
    require(file && dataset && ndim && dims, "null argument");
    Dataset d(file, dataset);
    *ndim = d.ndim;
    for (int i = 0; i < 3; ++i) dims[i] = i < d.ndim ? (long)d.dims[i] : 0;
  });
}

int tc_h5_read_projections(const char* file, const char* dataset, long begin,
                           long end, float* h_out) {
  return guarded([&] { read_block(file, dataset, begin, end, 1, h_out); });
}

int tc_h5_read_slices(const char* file, const char* dataset, long begin,
                      long end, float* h_out) {
  return guarded([&] { read_block(file, dataset, begin, end, 0, h_out); });
}

int tc_h5_read_1d(const char* file, const char* dataset, float* h_out,
                  long n) {
  return guarded([&] {
    require(file && dataset && h_out, "null argument");
    Dataset d(file, dataset);
    require(d.ndim == 1, std::string("dataset ") + dataset + " is not 1D");
    d.require_float(dataset);
    require((long)d.dims[0] == n, std::string("dataset ") + dataset +
                                      " has " + std::to_string(d.dims[0]) +
                                      " values, not " + std::to_string(n));
    if (H5Dread(d.dset.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                h_out) < 0)
      throw H5Error(2, std::string("reading ") + dataset + " failed");
  });
}

int tc_h5_write(const char* file, const char* dataset, const float* h_data,
                int ndim, const long* dims_in, int truncate) {
  return guarded([&] {
    require(file && dataset && h_data && dims_in, "null argument");
    require(ndim >= 1 && ndim <= 3, "rank must be 1, 2 or 3");
    hsize_t dims[3] = {1, 1, 1};
    for (int i = 0; i < ndim; ++i) {
      require(dims_in[i] >= 1, "empty array");
      dims[i] = (hsize_t)dims_in[i];
    }
    Handle f(truncate ? H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT,
                                  H5P_DEFAULT)
                      : H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT),
             H5Fclose, std::string("cannot open HDF5 file ") + file);
    Handle sp(H5Screate_simple(ndim, dims, nullptr), H5Sclose,
              "cannot create dataspace");
    Handle ds(H5Dcreate2(f.id, dataset, H5T_NATIVE_FLOAT, sp.id, H5P_DEFAULT,
                         H5P_DEFAULT, H5P_DEFAULT),
              H5Dclose, std::string("cannot create dataset ") + dataset);
    if (H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                 h_data) < 0)
      throw H5Error(2, std::string("writing ") + dataset + " failed");
  });
}

const char* tc_h5_last_error(void) { return g_err.c_str(); }

}  // extern "C"
