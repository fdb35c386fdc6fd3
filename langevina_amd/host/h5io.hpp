// h5io.hpp -- HDF5 output in the reference's layout:
//   * ntuple<rank> (ModelG/ntuple.h:49-273): a dataset [T, n1, ..., nrank]
//     of H5T_NATIVE_DOUBLE with an unlimited first dimension, chunk
//     {2, n1, ..}, extended by one row per fill();
//   * the o4fields restart dataset [NZ][NY][NX][10] (ModelA.h:356-402).
// libhdf5 (1.10.x) is opened with dlopen by absolute path, so that the
// library directory does not enter this program's search path.
#pragma once

#include <dlfcn.h>
#include <hdf5.h>

#include <array>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace lgv_host {

// H5F_ACC_* expand to calls of H5check()/H5open(), which are reached here
// only through dlsym; these are their values (HDF5 1.10 H5Fpublic.h:50-52).
constexpr unsigned kAccRdonly = 0x0000u, kAccRdwr = 0x0001u, kAccTrunc = 0x0002u;

struct H5Api {
  void *lib = nullptr;
  decltype(&H5open) open_;
  decltype(&H5Fcreate) Fcreate;
  decltype(&H5Fopen) Fopen;
  decltype(&H5Fclose) Fclose;
  decltype(&H5Screate_simple) Screate_simple;
  decltype(&H5Sclose) Sclose;
  decltype(&H5Sselect_hyperslab) Sselect_hyperslab;
  decltype(&H5Sget_simple_extent_dims) Sget_simple_extent_dims;
  decltype(&H5Pcreate) Pcreate;
  decltype(&H5Pset_chunk) Pset_chunk;
  decltype(&H5Pclose) Pclose;
  decltype(&H5Dcreate2) Dcreate2;
  decltype(&H5Dopen2) Dopen2;
  decltype(&H5Dclose) Dclose;
  decltype(&H5Dget_space) Dget_space;
  decltype(&H5Dset_extent) Dset_extent;
  decltype(&H5Dwrite) Dwrite;
  decltype(&H5Dread) Dread;
  decltype(&H5Lexists) Lexists;
  decltype(&H5Ldelete) Ldelete;
  decltype(&H5Sget_simple_extent_ndims) Sget_simple_extent_ndims;
  hid_t native_double = -1, dataset_create = -1;

  template <typename F>
  void sym(F &f, const char *name) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    if (!f) throw std::runtime_error(std::string("libhdf5: missing symbol ") + name);
  }

  static H5Api &get() {
    static H5Api api;
    if (!api.lib) api.load();
    return api;
  }

  void load() {
    const char *path = getenv("LGV_HDF5_LIB");
    const char *cands[] = {path, "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
                           "libhdf5.so"};
    for (const char *c : cands) {
      if (!c) continue;
      lib = dlopen(c, RTLD_NOW | RTLD_LOCAL);
      if (lib) break;
    }
    if (!lib) throw std::runtime_error("cannot dlopen libhdf5 (set LGV_HDF5_LIB)");
    sym(open_, "H5open");
    sym(Fcreate, "H5Fcreate");
    sym(Fopen, "H5Fopen");
    sym(Fclose, "H5Fclose");
    sym(Screate_simple, "H5Screate_simple");
    sym(Sclose, "H5Sclose");
    sym(Sselect_hyperslab, "H5Sselect_hyperslab");
    sym(Sget_simple_extent_dims, "H5Sget_simple_extent_dims");
    sym(Pcreate, "H5Pcreate");
    sym(Pset_chunk, "H5Pset_chunk");
    sym(Pclose, "H5Pclose");
    sym(Dcreate2, "H5Dcreate2");
    sym(Dopen2, "H5Dopen2");
    sym(Dclose, "H5Dclose");
    sym(Dget_space, "H5Dget_space");
    sym(Dset_extent, "H5Dset_extent");
    sym(Dwrite, "H5Dwrite");
    sym(Dread, "H5Dread");
    sym(Lexists, "H5Lexists");
    sym(Ldelete, "H5Ldelete");
    sym(Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
    if (open_() < 0) throw std::runtime_error("H5open failed");
    hid_t *nd = reinterpret_cast<hid_t *>(dlsym(lib, "H5T_NATIVE_DOUBLE_g"));
    hid_t *dc = reinterpret_cast<hid_t *>(dlsym(lib, "H5P_CLS_DATASET_CREATE_ID_g"));
    if (!nd || !dc) throw std::runtime_error("libhdf5: missing type/class ids");
    native_double = *nd;
    dataset_create = *dc;
  }
};

// ntuple<rank>, ModelG/ntuple.h:49-273
template <std::size_t rank>
class Ntuple {
 public:
  std::vector<double> row;

  Ntuple(const std::array<size_t, rank> &n, const std::string &name, hid_t file) {
    H5Api &h = H5Api::get();
    for (size_t i = 0; i < rank; i++) N_[i] = n[i];
    if (h.Lexists(file, name.c_str(), H5P_DEFAULT) <= 0) {
      pack(dims_, 0);
      hsize_t maxdims[rank + 1];
      pack(maxdims, H5S_UNLIMITED);
      space_ = h.Screate_simple(rank + 1, dims_, maxdims);
      hid_t cp = h.Pcreate(h.dataset_create);
      hsize_t chunk[rank + 1];
      pack(chunk, 2);
      h.Pset_chunk(cp, rank + 1, chunk);
      dset_ = h.Dcreate2(file, name.c_str(), h.native_double, space_, H5P_DEFAULT, cp,
                         H5P_DEFAULT);
      h.Pclose(cp);
    } else {
      dset_ = h.Dopen2(file, name.c_str(), H5P_DEFAULT);
      space_ = h.Dget_space(dset_);
      hsize_t maxdims[rank + 1];
      h.Sget_simple_extent_dims(space_, dims_, maxdims);
      for (size_t i = 0; i < rank; i++)
        if (N_[i] != dims_[i + 1])
          throw std::runtime_error("ntuple::ntuple Error dataset dimensions do not conform");
    }
    if (dset_ < 0) throw std::runtime_error("cannot create/open dataset " + name);
    size_t count = 1;
    for (size_t i = 0; i < rank; i++) count *= N_[i];
    row.assign(count, 0.);
    hsize_t line[rank + 1];
    pack(line, 1);
    mem_ = h.Screate_simple(rank + 1, line, nullptr);
  }
  // Opens an existing dataset [T, n1, .., nrank] for reading
  // (ntuple(name, file), ModelG/ntuple.h:108-110); row sizes come from the file.
  Ntuple(const std::string &name, hid_t file) {
    H5Api &h = H5Api::get();
    if (h.Lexists(file, name.c_str(), H5P_DEFAULT) <= 0)
      throw std::runtime_error("dataset " + name + " does not exist");
    dset_ = h.Dopen2(file, name.c_str(), H5P_DEFAULT);
    if (dset_ < 0) throw std::runtime_error("cannot open dataset " + name);
    space_ = h.Dget_space(dset_);
    if (h.Sget_simple_extent_ndims(space_) != (int)rank + 1)
      throw std::runtime_error("dataset " + name + " has rank " +
                               std::to_string(h.Sget_simple_extent_ndims(space_)) +
                               ", expected " + std::to_string(rank + 1));
    hsize_t maxdims[rank + 1];
    h.Sget_simple_extent_dims(space_, dims_, maxdims);
    size_t count = 1;
    for (size_t i = 0; i < rank; i++) {
      N_[i] = dims_[i + 1];
      count *= N_[i];
    }
    row.assign(count, 0.);
    hsize_t line[rank + 1];
    pack(line, 1);
    mem_ = h.Screate_simple(rank + 1, line, nullptr);
  }
  ~Ntuple() {
    H5Api &h = H5Api::get();
    h.Sclose(mem_);
    h.Dclose(dset_);
    h.Sclose(space_);
  }
  Ntuple(const Ntuple &) = delete;
  Ntuple &operator=(const Ntuple &) = delete;

  void fill() {
    H5Api &h = H5Api::get();
    dims_[0]++;
    h.Dset_extent(dset_, dims_);
    hid_t fs = h.Dget_space(dset_);
    hsize_t offset[rank + 1] = {};
    offset[0] = dims_[0] - 1;
    hsize_t count[rank + 1];
    pack(count, 1);
    h.Sselect_hyperslab(fs, H5S_SELECT_SET, offset, nullptr, count, nullptr);
    h.Dwrite(dset_, h.native_double, mem_, fs, H5P_DEFAULT, row.data());
    h.Sclose(fs);
  }
  size_t nrows() const { return (size_t)dims_[0]; }
  std::array<size_t, rank> getN() const {
    std::array<size_t, rank> n;
    for (size_t i = 0; i < rank; i++) n[i] = (size_t)N_[i];
    return n;
  }
  // row-major offset of an element of one row (ntuple::at, ntuple.h:98-102)
  size_t at(const std::array<size_t, rank> &idx) const {
    size_t k = 0;
    for (size_t i = 0; i < rank; i++) k = k * N_[i] + idx[i];
    return k;
  }
  double &Row(const std::array<size_t, rank> &idx) { return row[at(idx)]; }
  // row irow of the file into row (ntuple::readrow, ntuple.h:76-96)
  void readrow(size_t irow) {
    H5Api &h = H5Api::get();
    if (irow >= nrows()) throw std::runtime_error("ntuple::readrow: row out of range");
    hid_t fs = h.Dget_space(dset_);
    hsize_t offset[rank + 1] = {};
    offset[0] = irow;
    hsize_t count[rank + 1];
    pack(count, 1);
    h.Sselect_hyperslab(fs, H5S_SELECT_SET, offset, nullptr, count, nullptr);
    herr_t e = h.Dread(dset_, h.native_double, mem_, fs, H5P_DEFAULT, row.data());
    h.Sclose(fs);
    if (e < 0) throw std::runtime_error("ntuple::readrow: read failed");
  }

 private:
  hsize_t N_[rank];
  hsize_t dims_[rank + 1];
  hid_t space_ = -1, dset_ = -1, mem_ = -1;
  void pack(hsize_t *d, hsize_t d0) const {
    d[0] = d0;
    for (size_t i = 0; i < rank; i++) d[i + 1] = N_[i];
  }
};

class H5File {
 public:
  hid_t id = -1;
  // FILE_MODE_WRITE truncates; append opens an existing file read-write or
  // creates it (measurer_output.cxx:10-31)
  H5File(const std::string &name, bool append) {
    H5Api &h = H5Api::get();
    if (append) {
      FILE *f = fopen(name.c_str(), "rb");
      if (f) {
        fclose(f);
        id = h.Fopen(name.c_str(), kAccRdwr, H5P_DEFAULT);
      }
    }
    if (id < 0) id = h.Fcreate(name.c_str(), kAccTrunc, H5P_DEFAULT, H5P_DEFAULT);
    if (id < 0) throw std::runtime_error("cannot open " + name);
  }
  // an existing file, read-only or read-write; throws if it cannot be opened
  struct Existing {};
  H5File(const std::string &name, Existing, bool rw) {
    H5Api &h = H5Api::get();
    FILE *f = fopen(name.c_str(), "rb");
    if (f) {
      fclose(f);
      id = h.Fopen(name.c_str(), rw ? kAccRdwr : kAccRdonly, H5P_DEFAULT);
    }
    if (id < 0) throw std::runtime_error("Unable to open " + name);
  }
  ~H5File() {
    if (id >= 0) H5Api::get().Fclose(id);
  }
  H5File(const H5File &) = delete;
  H5File &operator=(const H5File &) = delete;
};

// ModelA::write / read: dataset o4fields [N][N][N][10] of the whole field
inline void write_o4fields(const std::string &fname, int N, const double *u) {
  H5Api &h = H5Api::get();
  H5File f(fname, false);
  hsize_t dims[4] = {(hsize_t)N, (hsize_t)N, (hsize_t)N, 10};
  hid_t sp = h.Screate_simple(4, dims, nullptr);
  hid_t ds = h.Dcreate2(f.id, "o4fields", h.native_double, sp, H5P_DEFAULT, H5P_DEFAULT,
                        H5P_DEFAULT);
  if (ds < 0) throw std::runtime_error("cannot create o4fields in " + fname);
  h.Dwrite(ds, h.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, u);
  h.Dclose(ds);
  h.Sclose(sp);
}

inline void read_o4fields(const std::string &fname, int N, double *u) {
  H5Api &h = H5Api::get();
  FILE *probe = fopen(fname.c_str(), "rb");
  if (!probe)
    throw std::string("Unable to open file in restart mode with filename = ") + fname;
  fclose(probe);
  hid_t fid = h.Fopen(fname.c_str(), kAccRdonly, H5P_DEFAULT);
  if (fid < 0) throw std::string("Unable to open file in restart mode with filename = ") + fname;
  hid_t ds = h.Dopen2(fid, "o4fields", H5P_DEFAULT);
  if (ds < 0) {
    h.Fclose(fid);
    throw std::string("no o4fields dataset in ") + fname;
  }
  hid_t sp = h.Dget_space(ds);
  hsize_t dims[4] = {0, 0, 0, 0}, maxd[4];
  h.Sget_simple_extent_dims(sp, dims, maxd);
  h.Sclose(sp);
  if (dims[0] != (hsize_t)N || dims[1] != (hsize_t)N || dims[2] != (hsize_t)N || dims[3] != 10) {
    h.Dclose(ds);
    h.Fclose(fid);
    throw std::string("o4fields in ") + fname + " does not match the lattice";
  }
  h.Dread(ds, h.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, u);
  h.Dclose(ds);
  h.Fclose(fid);
}

}  // namespace lgv_host
