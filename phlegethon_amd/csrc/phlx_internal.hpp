// phlx_internal.hpp -- shared declarations between the kernels and the C-ABI.
#pragma once

#include "phlx_device.hpp"
#include "phlx_helm.hpp"

namespace phlx {

// Launchers of one template instance (space dims, advected scalars, MHD).
struct KernelTable {
  void (*flux)(const Grid &, int, hipStream_t);
  void (*corner)(const Grid &, hipStream_t);
  void (*faceb)(const Grid &, int, double, double, double, hipStream_t);
  void (*update)(const Grid &, int, int, double, double, double, hipStream_t);
  void (*cfl)(const Grid &, hipStream_t);
  void (*init)(const Grid &, hipStream_t);
  void (*divb)(const Grid &, hipStream_t);
};

bool select_kernels(int sd, int nas, bool mhd, bool helm, KernelTable &t);
void launch_step_begin(Scalars *sc, hipStream_t st);
void launch_step_end(const Grid &g, int advance, hipStream_t st);
void launch_reduce_to(const Grid &g, double *out, hipStream_t st);
int cell_grid_blocks(const Grid &g);
void launch_stream_copy(const void *src, void *dst, long long n16, hipStream_t st);

// Helmholtz EoS (phlx_helm.hip)
void launch_helm_pack(long long np, const double *f9, const double *c4, double *tab,
                      hipStream_t st);
void launch_helm_full(const HelmDev &h, long long n, const double *rho, const double *T,
                      const double *abar, const double *zbar, double *P, double *E, double *cs,
                      double *cv, hipStream_t st);
void launch_helm_rhoe(const HelmDev &h, long long n, const double *rho, const double *e,
                      const double *abar, const double *zbar, double *T, double *P, double *cs,
                      int ret, int *iters, hipStream_t st);
void launch_helm_rhoP(const HelmDev &h, long long n, const double *rho, const double *p,
                      const double *abar, const double *zbar, double *T, double *E, double *cs,
                      int ret, int *iters, hipStream_t st);

enum { KID_FLUX1 = 0, KID_FLUX2, KID_FLUX3, KID_CORNER, KID_FACEB, KID_UPDATE, KID_MISC, KID_COUNT };

}  // namespace phlx
