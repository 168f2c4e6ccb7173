// pipio.cpp -- libpipio.so (declared in include/pipio.h): HDF5 snapshots in
// the dataset layout of the reference's output and restart routines
// (IO_rot.f90), for a single process holding the whole padded domain.
//
// The reference writes with the Fortran HDF5 API, which reverses the order
// of dimensions: the Fortran array U(ix,jx,kx) appears in the file as a
// (kx, jx, ix) dataspace in C order, with the same bytes.  Here the arrays
// are already in that memory order, so every 3-D field is written and read
// with one contiguous transfer.
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/pipio.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

// Variable names of set_initial_out (IO_rot.f90:362-374).
const char* const kNamesM[9] = {"ro_p", "mx_p", "my_p", "mz_p", "en_p", "bx", "by", "bz", "ps"};
const char* const kNamesH[5] = {"ro_n", "mx_n", "my_n", "mz_n", "en_n"};

std::string snapshot_path(const char* dir, int nout) {
  // trim(outdir) // 't' // tno // '.h5' (IO_rot.f90:114-116), with the
  // directory separator supplied when the directory string lacks it
  std::string p = dir ? dir : "";
  if (!p.empty() && p.back() != '/') p += '/';
  char name[32];
  snprintf(name, sizeof(name), "t%04d.h5", nout);
  return p + name;
}

struct File {
  hid_t id = -1;
  ~File() {
    if (id >= 0) H5Fclose(id);
  }
};

int write_ds(hid_t file, const char* name, int rank, const hsize_t* dims, const double* data) {
  hid_t space = (rank == 0) ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, nullptr);
  if (space < 0) return fail(std::string("dataspace for ") + name);
  hid_t ds = H5Dcreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) {
    H5Sclose(space);
    return fail(std::string("cannot create dataset ") + name);
  }
  herr_t e = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  H5Sclose(space);
  return (e < 0) ? fail(std::string("cannot write dataset ") + name) : 0;
}

int read_ds(hid_t file, const char* name, int rank, const hsize_t* dims, double* out) {
  hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
  if (ds < 0) return fail(std::string("no dataset ") + name);
  hid_t space = H5Dget_space(ds);
  const int r = H5Sget_simple_extent_ndims(space);
  bool ok = (r == rank);
  if (ok && rank > 0) {
    hsize_t d[3] = {0, 0, 0};
    H5Sget_simple_extent_dims(space, d, nullptr);
    for (int k = 0; k < rank; k++) ok = ok && (d[k] == dims[k]);
  }
  H5Sclose(space);
  if (!ok) {
    H5Dclose(ds);
    return fail(std::string("dataset ") + name + " has another shape than the layout");
  }
  herr_t e = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
  H5Dclose(ds);
  return (e < 0) ? fail(std::string("cannot read dataset ") + name) : 0;
}

int check_layout(const pipio_layout* L) {
  if (!L) return fail("null layout");
  if (L->ix < 1 || L->jx < 1 || L->kx < 1) return fail("sizes must be >= 1");
  if (L->nvar_m != 8 && L->nvar_m != 9) return fail("nvar_m must be 8 or 9");
  return 0;
}

#define TRY(x)           \
  do {                   \
    if ((x) != 0) return -1; \
  } while (0)

}  // namespace

extern "C" {

int pipio_write(const char* outdir, int nout, const pipio_layout* L, const pipio_fields* f) {
  TRY(check_layout(L));
  if (!f || !f->U_h || !f->U_m || !f->x || !f->y || !f->z || !f->dx || !f->dy || !f->dz)
    return fail("null state or coordinate array");
  if (nout < 0 || nout > 9999) return fail("nout must be in 0..9999 (t####.h5)");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::string path = snapshot_path(outdir, nout);
  File file;
  file.id = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file.id < 0) return fail("cannot create " + path);
  const hsize_t nx = L->ix, ny = L->jx, nz = L->kx;
  const hsize_t d3[3] = {nz, ny, nx};
  const size_t plane = (size_t)nx * ny * nz;
  // save_coordinates, write_time (IO_rot.f90:182-206)
  TRY(write_ds(file.id, "xgrid", 1, &nx, f->x));
  TRY(write_ds(file.id, "ygrid", 1, &ny, f->y));
  TRY(write_ds(file.id, "zgrid", 1, &nz, f->z));
  TRY(write_ds(file.id, "dx", 1, &nx, f->dx));
  TRY(write_ds(file.id, "dy", 1, &ny, f->dy));
  TRY(write_ds(file.id, "dz", 1, &nz, f->dz));
  TRY(write_ds(file.id, "time", 0, nullptr, &f->time));
  // def_varfiles (IO_rot.f90:231-259) for the two-fluid run (flag_pip = 1,
  // no resistivity, gravity or viscosity); a name the reference would
  // create twice keeps its first dataset
  std::vector<std::string> done;
  auto put = [&](const char* name, const double* data) -> int {
    for (const auto& s : done)
      if (s == name) return 0;
    done.emplace_back(name);
    return write_ds(file.id, name, 3, d3, data);
  };
  if (f->ac && L->ac_sav == 0) TRY(put("ac", f->ac));
  if (f->xi && L->xi_sav == 0) TRY(put("xi", f->xi));
  if (L->flag_ir == 1) {
    if (f->gm_ion && L->ion_sav == 0) TRY(put("ion", f->gm_ion));
    if (f->gm_rec && L->rec_sav == 0) TRY(put("rec", f->gm_rec));
  }
  if (L->flag_col == 1 && f->ac && L->col_sav == 0) TRY(put("col", f->ac));
  if (L->flag_ir_type == 0 && L->flag_ir != 0 && f->arb_heat && L->heat_sav == 0) TRY(put("aheat", f->arb_heat));
  // save_varfiles (IO_rot.f90:284-330)
  for (int i = 0; i < L->nvar_m; i++)
    if (i < 8 || (L->flag_divb == 1 && L->ps_sav == 0)) TRY(put(kNamesM[i], f->U_m + (size_t)i * plane));
  if (L->flag_ir >= 1) {
    if (f->gm_ion && L->ion_sav == 0) TRY(put("ion", f->gm_ion));
    if (f->gm_rec && L->rec_sav == 0) TRY(put("rec", f->gm_rec));
  }
  for (int i = 0; i < 5; i++) TRY(put(kNamesH[i], f->U_h + (size_t)i * plane));
  if (H5Fflush(file.id, H5F_SCOPE_LOCAL) < 0) return fail("cannot flush " + path);
  return 0;
}

int pipio_read(const char* indir, int nout, const pipio_layout* L, double* time, double* x, double* y, double* z,
               double* dx, double* dy, double* dz, double* U_h, double* U_m, double* ac, double* gm_ion,
               double* gm_rec, double* arb_heat) {
  TRY(check_layout(L));
  if (!U_h || !U_m || !time) return fail("null state or time output");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const std::string path = snapshot_path(indir, nout);
  File file;
  file.id = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file.id < 0) return fail("cannot open " + path);
  const hsize_t nx = L->ix, ny = L->jx, nz = L->kx;
  const hsize_t d3[3] = {nz, ny, nx};
  const size_t plane = (size_t)nx * ny * nz;
  // reread_coordinate (IO_rot.f90:524-533)
  if (x) TRY(read_ds(file.id, "xgrid", 1, &nx, x));
  if (y) TRY(read_ds(file.id, "ygrid", 1, &ny, y));
  if (z) TRY(read_ds(file.id, "zgrid", 1, &nz, z));
  if (dx) TRY(read_ds(file.id, "dx", 1, &nx, dx));
  if (dy) TRY(read_ds(file.id, "dy", 1, &ny, dy));
  if (dz) TRY(read_ds(file.id, "dz", 1, &nz, dz));
  // reread_variables (IO_rot.f90:557-593): the plasma's nvar_m planes, the
  // neutrals, then the coefficient fields and time
  for (int i = 0; i < L->nvar_m; i++) TRY(read_ds(file.id, kNamesM[i], 3, d3, U_m + (size_t)i * plane));
  for (int i = 0; i < 5; i++) TRY(read_ds(file.id, kNamesH[i], 3, d3, U_h + (size_t)i * plane));
  if (ac) TRY(read_ds(file.id, "ac", 3, d3, ac));
  if (gm_ion) TRY(read_ds(file.id, "ion", 3, d3, gm_ion));
  if (gm_rec) TRY(read_ds(file.id, "rec", 3, d3, gm_rec));
  if (arb_heat) TRY(read_ds(file.id, "aheat", 3, d3, arb_heat));
  TRY(read_ds(file.id, "time", 0, nullptr, time));
  return 0;
}

const char* pipio_last_error(void) { return g_err.c_str(); }

}  // extern "C"
