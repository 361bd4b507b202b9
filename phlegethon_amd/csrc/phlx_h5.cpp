// phlx_h5.cpp -- grid output in the reference's HDF5 layout (write_output,
// source/source.F90:2899-3169, and its helpers hdf5_annotate_rp/_ip/_string
// :5571-5686, hdf5_write_array/_ndarray :5100-5450 with compression factor
// 1).  Host code only; built into libphlx_h5.so against HDF5 1.10.
#include <hdf5.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/phlx_h5.h"

namespace {

struct H5Err {
  bool bad = false;
  void check(long long rc, const char *what) {
    if (rc < 0 && !bad) {
      std::fprintf(stderr, "phlx_h5: %s failed\n", what);
      bad = true;
    }
  }
};

// scalar dataset (hdf5_annotate_rp / _ip write datasets, not attributes)
void scalar_f64(hid_t g, const char *key, double v, H5Err &e) {
  hid_t s = H5Screate(H5S_SCALAR);
  hid_t d = H5Dcreate2(g, key, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  e.check(d, key);
  if (d >= 0) e.check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), key);
  if (d >= 0) H5Dclose(d);
  H5Sclose(s);
}

void scalar_i32(hid_t g, const char *key, int v, H5Err &e) {
  hid_t s = H5Screate(H5S_SCALAR);
  hid_t d = H5Dcreate2(g, key, H5T_STD_I32LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  e.check(d, key);
  if (d >= 0) e.check(H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), key);
  if (d >= 0) H5Dclose(d);
  H5Sclose(s);
}

// fixed-length string attribute of the trimmed length (hdf5_annotate_string)
void attr_str(hid_t g, const char *key, const char *val, H5Err &e) {
  const std::string v(val);
  hid_t t = H5Tcopy(H5T_FORTRAN_S1);
  H5Tset_size(t, v.size());
  hid_t s = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(g, key, t, s, H5P_DEFAULT, H5P_DEFAULT);
  e.check(a, key);
  if (a >= 0) e.check(H5Awrite(a, t, v.data()), key);
  if (a >= 0) H5Aclose(a);
  H5Sclose(s);
  H5Tclose(t);
}

// F64 dataset of C dims (n3, n2, n1[, nc]) from a buffer in that order
void array_f64(hid_t g, const char *key, const std::vector<double> &buf, hsize_t n3,
               hsize_t n2, hsize_t n1, hsize_t nc, H5Err &e) {
  const hsize_t dims[4] = {n3, n2, n1, nc};
  hid_t s = H5Screate_simple(nc ? 4 : 3, dims, nullptr);
  hid_t d = H5Dcreate2(g, key, H5T_IEEE_F64LE, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  e.check(d, key);
  if (d >= 0)
    e.check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), key);
  if (d >= 0) H5Dclose(d);
  H5Sclose(s);
}

}  // namespace

extern "C" int phlx_h5_write_grid(const char *dir, const phlx_config *cfg, int64_t step,
                                  double time, double dt, const double *prim,
                                  const double *temp, const double *b_cc,
                                  const double *grav_x2, const double *phi_x2,
                                  const double *edot_x2, const double *kappa) {
  if (!dir || !cfg || !prim || step < 0) return -1;
  const int sd = cfg->sdims;
  const hsize_t n1 = cfg->nx1, n2 = cfg->nx2, n3 = sd == 3 ? cfg->nx3 : 1;
  const hsize_t N = n1 * n2 * n3;
  const int nv = sd + 2 + cfg->nas;
  const bool mhd = cfg->flux == PHLX_HLLD_FLUX || cfg->flux == PHLX_LHLLD_FLUX;
  const bool helm = cfg->eos == PHLX_EOS_HELMHOLTZ;
  const bool pig = cfg->eos == PHLX_EOS_PIG;
  // ADVECT_YE_IABAR: the two advected scalars of a tabulated EoS (Helmholtz
  // always carries them, PIG may)
  const bool yeiabar = cfg->eos != PHLX_EOS_IDEAL && cfg->nas >= 2;
  char name[64];
  std::snprintf(name, sizeof(name), "/grid_n%05lld.h5", (long long)step);
  const std::string path = std::string(dir) + name;
  H5Err e;
  hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) {
    std::fprintf(stderr, "phlx_h5: cannot create %s\n", path.c_str());
    return -1;
  }
  hid_t g = H5Gcreate2(f, "grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  e.check(g, "grid");
  if (g >= 0) {
    scalar_f64(g, "time", time, e);
    scalar_f64(g, "dt", dt, e);
    scalar_i32(g, "step", (int)step, e);
    const double dx[3] = {(cfg->x1u - cfg->x1l) / cfg->nx1, (cfg->x2u - cfg->x2l) / cfg->nx2,
                          sd == 3 ? (cfg->x3u - cfg->x3l) / cfg->nx3 : 1.0};
    if (step == 0) {
      scalar_f64(g, "gamma_ad", cfg->gamma, e);
      scalar_f64(g, "mu", cfg->mu, e);
      scalar_f64(g, "x1l", cfg->x1l, e);
      scalar_f64(g, "x1u", cfg->x1u, e);
      scalar_f64(g, "x2l", cfg->x2l, e);
      scalar_f64(g, "x2u", cfg->x2u, e);
      scalar_f64(g, "x3l", cfg->x3l, e);
      scalar_f64(g, "x3u", cfg->x3u, e);
      scalar_i32(g, "sdims", sd, e);
      scalar_i32(g, "nx1", cfg->nx1, e);
      scalar_i32(g, "nx2", cfg->nx2, e);
      scalar_i32(g, "nx3", cfg->nx3, e);
      attr_str(g, "advect_yeiabar", yeiabar ? "true" : "false", e);
      attr_str(g, "advect_species", "false", e);
      attr_str(g, "use_prad", "false", e);
      attr_str(g, "use_helmholtz", helm ? "true" : "false", e);
      if (helm) attr_str(g, "use_coulomb_corrections", "false", e);
      attr_str(g, "use_pig", pig ? "true" : "false", e);
      attr_str(g, "use_gravity_solver", "false", e);
      attr_str(g, "use_monopole_solver", "false", e);
      attr_str(g, "geometry-type", "cartesian", e);
      attr_str(g, "use_internal_boundaries", "false", e);
      // cell centres x_l + (i - 1/2) dx with 1-based i (:16296-16299)
      std::vector<double> buf(N * sd);
      for (hsize_t k = 0; k < n3; k++)
        for (hsize_t j = 0; j < n2; j++)
          for (hsize_t i = 0; i < n1; i++) {
            const hsize_t c = (k * n2 + j) * n1 + i;
            const double idx[3] = {(double)(i + 1), (double)(j + 1), (double)(k + 1)};
            const double lo[3] = {cfg->x1l, cfg->x2l, cfg->x3l};
            for (int d = 0; d < sd; d++) buf[c * sd + d] = lo[d] + (idx[d] - 0.5) * dx[d];
          }
      attr_str(g, "update_kappa", "false", e);
      array_f64(g, "coords", buf, n3, n2, n1, sd, e);
      if (kappa) {
        // THERMAL_DIFFUSION_STS without USE_TIMMES_KAPPA (:3070-3076)
        std::vector<double> s(kappa, kappa + N);
        array_f64(g, "kappa", s, n3, n2, n1, 0, e);
      }
      if (cfg->use_gravity && grav_x2) {
        // plane-parallel gravity: grav(2, i, j, k) = g(j), other components 0
        for (hsize_t c = 0; c < N * sd; c++) buf[c] = 0.0;
        for (hsize_t c = 0; c < N; c++) buf[c * sd + 1] = grav_x2[(c / n1) % n2];
        array_f64(g, "grav", buf, n3, n2, n1, sd, e);
        if (cfg->evolve_etot && phi_x2) {
          std::vector<double> s(N);
          for (hsize_t c = 0; c < N; c++) s[c] = phi_x2[(c / n1) % n2];
          array_f64(g, "phi_cc", s, n3, n2, n1, 0, e);
        }
      }
      if (cfg->use_edot && edot_x2) {
        std::vector<double> s(N);
        for (hsize_t c = 0; c < N; c++) s[c] = edot_x2[(c / n1) % n2];
        array_f64(g, "edot", s, n3, n2, n1, 0, e);
      }
      // vol = 1 / ivol with ivol = 1 / (dx1 dx2 [dx3]) (:5799-5806, :5859)
      double cell = dx[0] * dx[1];
      if (sd == 3) cell = dx[0] * dx[1] * dx[2];
      const double ivol = 1.0 / cell;
      std::vector<double> s(N, 1.0 / ivol);
      array_f64(g, "vol", s, n3, n2, n1, 0, e);
    }
    // prim(1:nvars, i, j, k): the variable index fastest
    std::vector<double> buf(N * nv);
    for (hsize_t c = 0; c < N; c++)
      for (int v = 0; v < nv; v++) buf[c * nv + v] = prim[(hsize_t)v * N + c];
    array_f64(g, "prim", buf, n3, n2, n1, nv, e);
    std::vector<double> t(N, 0.0);
    if (temp)
      for (hsize_t c = 0; c < N; c++) t[c] = temp[c];
    array_f64(g, "temp", t, n3, n2, n1, 0, e);
    if (mhd && b_cc) {
      std::vector<double> b(N * sd);
      for (hsize_t c = 0; c < N; c++)
        for (int d = 0; d < sd; d++) b[c * sd + d] = b_cc[(hsize_t)d * N + c];
      array_f64(g, "bfield", b, n3, n2, n1, sd, e);
    }
    H5Gclose(g);
  }
  H5Fclose(f);
  return e.bad ? -1 : 0;
}
