/* phlx_h5.h -- grid output in the reference's HDF5 layout (libphlx_h5.so, a
 * host-only library beside libphlx.so, linked against HDF5 1.10).
 *
 * Replaces write_output (source/source.F90:2899-3169) for the configurations
 * libphlx runs (Cartesian uniform grids, no diffusion, species, networks or
 * gravity solvers): file <dir>/grid_nNNNNN.h5 (step, 5 digits), group
 * "grid" holding scalar datasets time, dt (F64) and step (I32); at step 0
 * also gamma_ad, mu, x1l..x3u (F64), sdims, nx1..nx3 (I32), the string
 * attributes of the build flags (use_helmholtz / use_pig from the EoS,
 * advect_yeiabar for a tabulated EoS with two advected scalars), and the
 * datasets coords (cell centres), kappa (THERMAL_DIFFUSION_STS), grav
 * (USE_GRAVITY), phi_cc (USE_GRAVITY + EVOLVE_ETOT), edot (USE_EDOT) and
 * vol; every step prim, temp and bfield (USE_MHD).  Arrays are stored as
 * the reference's Fortran writer leaves them: dims (nx3, nx2, nx1[, ncomp])
 * in C order, component fastest, F64 little endian.
 *
 * Inputs are host arrays in libphlx's SoA order (phlx_download_state):
 * prim [nv][nx3][nx2][nx1], temp [nx3][nx2][nx1], b_cc [3][nx3][nx2][nx1];
 * the static fields as x2 profiles (length nx2; NULL when the flag is off);
 * kappa on the interior cells [nx3][nx2][nx1] (NULL without thermal
 * diffusion).
 * Returns 0, or -1 with a message on stderr when HDF5 fails. */
#ifndef PHLX_H5_H
#define PHLX_H5_H

#include <stdint.h>

#include "phlx.h"

#ifdef __cplusplus
extern "C" {
#endif

int phlx_h5_write_grid(const char *dir, const phlx_config *cfg, int64_t step, double time,
                       double dt, const double *prim, const double *temp, const double *b_cc,
                       const double *grav_x2, const double *phi_x2, const double *edot_x2,
                       const double *kappa);

#ifdef __cplusplus
}
#endif

#endif
