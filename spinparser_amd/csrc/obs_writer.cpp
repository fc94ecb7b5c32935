// obs_writer.cpp -- HDF5 observable files in the reference's .obs layout
// (host-only; see include/spobs.h for the layout and the reference sections it
// reproduces).
#include <hdf5.h>

#include <string>
#include <vector>

#include "spobs.h"

namespace {

thread_local std::string g_err;

int err(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// RAII closers for HDF5 identifiers
struct H {
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  H(const H &) = delete;
  H &operator=(const H &) = delete;
};

// float[3] array datasets of the meta group
int write_vec3(hid_t parent, const char *name, int ndim, const hsize_t *dims, const float *xyz) {
  const hsize_t three[1] = {3};
  H type(H5Tarray_create(H5T_NATIVE_FLOAT, 1, three), H5Tclose);
  H space(H5Screate_simple(ndim, dims, nullptr), H5Sclose);
  if (type.id < 0 || space.id < 0) return err(SP_OBS_EIO, std::string("cannot create the type of meta/") + name);
  H ds(H5Dcreate(parent, name, type.id, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (ds.id < 0) return err(SP_OBS_EIO, std::string("cannot create meta/") + name);
  if (H5Dwrite(ds.id, type.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, xyz) < 0)
    return err(SP_OBS_EIO, std::string("cannot write meta/") + name);
  return SP_OBS_OK;
}

// measurement groups of /<group>/data, in index order
std::vector<std::string> measurement_names(hid_t data) {
  std::vector<std::string> out;
  hsize_t n = 0;
  if (H5Gget_num_objs(data, &n) < 0) return out;
  for (hsize_t i = 0; i < n; ++i) {
    if (H5Gget_objtype_by_idx(data, i) != H5G_GROUP) continue;
    char name[64];
    if (H5Gget_objname_by_idx(data, i, name, sizeof name) < 0) continue;
    out.emplace_back(name);
  }
  return out;
}

}  // namespace

extern "C" {

const char *sp_obs_last_error(void) { return g_err.c_str(); }

int sp_obs_write_correlation(const char *path, const char *group, float cutoff, int n_bravais, const float *bravais,
                             int B, const float *basis, int Rb, const float *sites, const float *data) {
  if (!path || !group || !bravais || !basis || !sites || !data || n_bravais < 1 || B < 1 || Rb < 1)
    return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fis_hdf5(path) > 0 ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                               : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
         H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open observable file ") + path);
  H grp(H5Lexists(file.id, group, H5P_DEFAULT) > 0 ? H5Gopen(file.id, group, H5P_DEFAULT)
                                                    : H5Gcreate(file.id, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Gclose);
  if (grp.id < 0) return err(SP_OBS_EIO, std::string("could not open group ") + group);
  if (H5Lexists(grp.id, "meta", H5P_DEFAULT) <= 0) {
    H meta(H5Gcreate(grp.id, "meta", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (meta.id < 0) return err(SP_OBS_EIO, "could not create the meta group");
    const hsize_t dl[1] = {(hsize_t)n_bravais}, db[1] = {(hsize_t)B}, ds[2] = {(hsize_t)B, (hsize_t)Rb};
    int rc;
    if ((rc = write_vec3(meta.id, "latticeVectors", 1, dl, bravais)) != SP_OBS_OK) return rc;
    if ((rc = write_vec3(meta.id, "basis", 1, db, basis)) != SP_OBS_OK) return rc;
    if ((rc = write_vec3(meta.id, "sites", 2, ds, sites)) != SP_OBS_OK) return rc;
  }
  H dat(H5Lexists(grp.id, "data", H5P_DEFAULT) > 0 ? H5Gopen(grp.id, "data", H5P_DEFAULT)
                                                    : H5Gcreate(grp.id, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Gclose);
  if (dat.id < 0) return err(SP_OBS_EIO, "could not open the data group");
  // duplicate check on the cutoff attribute of every existing measurement
  const std::vector<std::string> names = measurement_names(dat.id);
  for (const std::string &nm : names) {
    H m(H5Gopen(dat.id, nm.c_str(), H5P_DEFAULT), H5Gclose);
    H a(H5Aopen(m.id, "cutoff", H5P_DEFAULT), H5Aclose);
    float c = 0.0f;
    if (a.id >= 0 && H5Aread(a.id, H5T_NATIVE_FLOAT, &c) >= 0 && c == cutoff) return SP_OBS_DUPLICATE;
  }
  const std::string name = "measurement_" + std::to_string(names.size());
  H m(H5Gcreate(dat.id, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (m.id < 0) return err(SP_OBS_EIO, "could not create " + name);
  const hsize_t one[1] = {1};
  H aspace(H5Screate_simple(1, one, nullptr), H5Sclose);
  H attr(H5Acreate(m.id, "cutoff", H5T_NATIVE_FLOAT, aspace.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (attr.id < 0 || H5Awrite(attr.id, H5T_NATIVE_FLOAT, &cutoff) < 0) return err(SP_OBS_EIO, "cannot write cutoff");
  const hsize_t dims[2] = {(hsize_t)B, (hsize_t)Rb};
  H space(H5Screate_simple(2, dims, nullptr), H5Sclose);
  H ds(H5Dcreate(m.id, "data", H5T_NATIVE_FLOAT, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (ds.id < 0 || H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return err(SP_OBS_EIO, "cannot write " + name + "/data");
  return SP_OBS_OK;
}

int sp_obs_count(const char *path, const char *group) {
  if (!path || !group) return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  if (H5Fis_hdf5(path) <= 0) return 0;
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open ") + path);
  if (H5Lexists(file.id, group, H5P_DEFAULT) <= 0) return 0;
  H grp(H5Gopen(file.id, group, H5P_DEFAULT), H5Gclose);
  if (H5Lexists(grp.id, "data", H5P_DEFAULT) <= 0) return 0;
  H dat(H5Gopen(grp.id, "data", H5P_DEFAULT), H5Gclose);
  return (int)measurement_names(dat.id).size();
}

int sp_obs_read_correlation(const char *path, const char *group, int k, float *cutoff, float *data, int n) {
  if (!path || !group || !cutoff || !data || k < 0 || n < 1) return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open ") + path);
  const std::string m = std::string(group) + "/data/measurement_" + std::to_string(k);
  H g(H5Gopen(file.id, m.c_str(), H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return err(SP_OBS_EIO, "no group " + m);
  H a(H5Aopen(g.id, "cutoff", H5P_DEFAULT), H5Aclose);
  if (a.id < 0 || H5Aread(a.id, H5T_NATIVE_FLOAT, cutoff) < 0) return err(SP_OBS_EIO, "no cutoff in " + m);
  H ds(H5Dopen(g.id, "data", H5P_DEFAULT), H5Dclose);
  if (ds.id < 0) return err(SP_OBS_EIO, "no data in " + m);
  H space(H5Dget_space(ds.id), H5Sclose);
  const hssize_t count = H5Sget_simple_extent_npoints(space.id);
  if (count != n) return err(SP_OBS_EARG, "size mismatch reading " + m);
  if (H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return err(SP_OBS_EIO, "read " + m);
  return SP_OBS_OK;
}

int sp_ckpt_write(const char *path, int append, float cutoff, int narrays, const char *const *names,
                  const long long *sizes, const float *const *data) {
  if (!path || narrays < 0 || (narrays > 0 && (!names || !sizes || !data))) return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  H file(append && H5Fis_hdf5(path) > 0 ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                                         : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
         H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open data file ") + path);
  const std::vector<std::string> groups = measurement_names(file.id);
  for (const std::string &nm : groups) {
    H g(H5Gopen(file.id, nm.c_str(), H5P_DEFAULT), H5Gclose);
    H a(H5Aopen(g.id, "cutoff", H5P_DEFAULT), H5Aclose);
    float c = 0.0f;
    if (a.id >= 0 && H5Aread(a.id, H5T_NATIVE_FLOAT, &c) >= 0 && c == cutoff) return SP_OBS_DUPLICATE_CKPT;
  }
  const int id = (int)groups.size();
  const std::string name = "checkpoint_" + std::to_string(id);
  H g(H5Gcreate(file.id, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return err(SP_OBS_EIO, "could not create " + name);
  const hsize_t one[1] = {1};
  H aspace(H5Screate_simple(1, one, nullptr), H5Sclose);
  H attr(H5Acreate(g.id, "cutoff", H5T_NATIVE_FLOAT, aspace.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (attr.id < 0 || H5Awrite(attr.id, H5T_NATIVE_FLOAT, &cutoff) < 0) return err(SP_OBS_EIO, "cannot write cutoff");
  auto dataset = [&](const char *nm, hsize_t n, const float *v) {
    H space(H5Screate_simple(1, &n, nullptr), H5Sclose);
    H ds(H5Dcreate(g.id, nm, H5T_NATIVE_FLOAT, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return ds.id >= 0 && H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
  };
  if (!dataset("cutoff", 1, &cutoff)) return err(SP_OBS_EIO, "cannot write " + name + "/cutoff");
  for (int i = 0; i < narrays; ++i)
    if (!names[i] || !data[i] || sizes[i] < 1 || !dataset(names[i], (hsize_t)sizes[i], data[i]))
      return err(SP_OBS_EIO, "cannot write " + name + "/" + (names[i] ? names[i] : "?"));
  return id;
}

int sp_ckpt_count(const char *path) {
  if (!path) return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  if (H5Fis_hdf5(path) <= 0) return 0;
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open data file ") + path);
  return (int)measurement_names(file.id).size();
}

int sp_ckpt_read(const char *path, int id, float *cutoff, int narrays, const char *const *names,
                 const long long *sizes, float *const *data) {
  if (!path || !cutoff || narrays < 0 || (narrays > 0 && (!names || !sizes || !data)))
    return err(SP_OBS_EARG, "invalid argument");
  H5Eset_auto(H5E_DEFAULT, nullptr, nullptr);
  H file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return err(SP_OBS_EIO, std::string("could not open data file ") + path);
  if (id < 0) id = (int)measurement_names(file.id).size() - 1;
  const std::string name = "checkpoint_" + std::to_string(id);
  H g(H5Gopen(file.id, name.c_str(), H5P_DEFAULT), H5Gclose);
  if (g.id < 0) return err(SP_OBS_EIO, "no group " + name);
  auto read = [&](const char *nm, long long n, float *v) {
    H ds(H5Dopen(g.id, nm, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) return false;
    H space(H5Dget_space(ds.id), H5Sclose);
    if (H5Sget_simple_extent_npoints(space.id) != n) return false;
    return H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
  };
  if (!read("cutoff", 1, cutoff)) return err(SP_OBS_EIO, "no cutoff in " + name);
  for (int i = 0; i < narrays; ++i)
    if (!names[i] || !data[i] || !read(names[i], sizes[i], data[i]))
      return err(SP_OBS_EIO, "cannot read " + name + "/" + (names[i] ? names[i] : "?"));
  return id;
}

}  // extern "C"
