// phlx_internal.hpp -- shared declarations between the kernels and the C-ABI.
#pragma once

#include "phlx_device.hpp"
#include "phlx_helm.hpp"
#include "phlx_mlp.hpp"

namespace phlx {

// x1 tile edge of the fused sweep: it runs where it divides the grid
constexpr int kSweepT = 16;

// Where k_faceb takes b0 and puts the new face field (its comment has the
// four cases); RK2 steps alternate the face-field buffers instead of
// copying the step-start field.
enum { FACEB_LATER = 0, FACEB_FIRST = 1, FACEB_RK2_1 = 2, FACEB_RK2_2 = 3 };

// Launchers of one template instance (space dims, advected scalars, MHD).
struct KernelTable {
  void (*flux)(const Grid &, int, hipStream_t);
  void (*corner)(const Grid &, hipStream_t);
  void (*faceb)(const Grid &, int, double, double, double, hipStream_t);
  void (*update)(const Grid &, int, int, double, double, double, int, hipStream_t);
  void (*update_guess)(const Grid &, int, int, double, double, double, int, hipStream_t);
  void (*cfl)(const Grid &, hipStream_t);
  void (*init)(const Grid &, int restart, hipStream_t);
  void (*divb)(const Grid &, hipStream_t);
  // mixed precision: refresh the FP32 sweep inputs from the FP64 state
  void (*pack)(const Grid &, hipStream_t);
  void (*prims)(const Grid &, hipStream_t);   // mixed: FP64 prim / eint / b_cc from U
  // fused sweep (3D, 2nd-order reconstructions, ideal / FASTEOS states;
  // null otherwise): one pass forms the stage's conserved state, the swept
  // updates finish it after CT
  void (*sweep)(const Grid &, int, int, double, double, double, hipStream_t);
  void (*update_swept)(const Grid &, int, int, double, double, double, int, hipStream_t);
  void (*update_swept_guess)(const Grid &, int, int, double, double, double, int, hipStream_t);
};

bool select_kernels(int sd, int nas, bool mhd, bool helm, bool fasteos, bool mixed, bool r5,
                    KernelTable &t);
// the four instance families (phlx_sel_*.hip)
bool select_fp64(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t);
bool select_fp64_r5(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t);
bool select_mixed(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t);
bool select_mixed_r5(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t);
void launch_step_begin(Scalars *sc, hipStream_t st);
// device state flags: flags = (flags & keep) | set
void launch_flags(Scalars *sc, int keep, int set, hipStream_t st);
void launch_step_end(const Grid &g, int advance, hipStream_t st);
void launch_reduce_to(const Grid &g, double *out, hipStream_t st);
int cell_grid_blocks(const Grid &g);
void launch_stream_copy(const void *src, void *dst, long long n16, hipStream_t st);
// USE_SHOCK_FLATTENING flags of the current primitive state (g.sf_flag)
void launch_shock_flags(const Grid &g, hipStream_t st);

// THERMAL_DIFFUSION_STS (phlx_diffusion.hip)
void launch_diff_t_from_p(const Grid &g, hipStream_t st);
void launch_diff_stage(const Grid &g, bool first, double mu, double nu, double c0, double hm,
                       double hg, hipStream_t st);
void launch_diff_temp(const Grid &g, bool helm, hipStream_t st);
void launch_diff_final(const Grid &g, bool helm, hipStream_t st);
void launch_parabolic(const Grid &g, double inv_dl2, hipStream_t st);

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

// T(rho, e) surrogate (phlx_mlp.hip)
void launch_mlp_data(const HelmDev &h, long long n, unsigned long long seed, const double box[8],
                     float *raw, float *label, float *wt, hipStream_t st);
void launch_mlp_normalize(const MlpDev &m, long long n, const float *raw, const float *label,
                          float *xn, float *tn, hipStream_t st);
int launch_mlp_grad(const float *theta, const float *w1q, const float *w3q, const float *xn,
                    const float *tn, const float *wt, long long n0, long long nb, float *gpart,
                    float *lpart, hipStream_t st);
void launch_mlp_adam(float *theta, float *m, float *v, const float *gpart, int nblk, int t,
                     float lr, float *w1q, float *w3q, unsigned short *wa, float *gout,
                     hipStream_t st);
void launch_mlp_quantize(const float *theta, float *w1q, float *w3q, unsigned short *wa,
                         hipStream_t st);
void launch_mlp_forward(const MlpDev &m, long long n, const float *raw, float *y, int *inbox,
                        hipStream_t st);
void launch_helm_rhoe_guess(const HelmDev &h, const MlpDev &m, int mode, long long n,
                            const double *rho, const double *e, const double *abar,
                            const double *zbar, double *T, double *P, double *cs, int ret,
                            int *iters, int *fb, hipStream_t st);

enum {
  KID_FLUX1 = 0, KID_FLUX2, KID_FLUX3, KID_CORNER, KID_FACEB, KID_UPDATE, KID_MISC, KID_DIFF,
  KID_SWEEP, KID_COUNT
};

}  // namespace phlx
