/*
 * oracle.h -- CPU FP64 restatement of Phlegethon's explicit finite-volume
 * update (source/source.F90 hydro_step :6847-13701, riemann :13969-16260,
 * compute_hyperbolic_dt :13707-13960, time_loop stepping :6271-6724).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for
 * the HIP path and the CPU baseline of the benchmark.  The product library
 * (phlegethon_amd/csrc) never links, loads or calls it.
 *
 * Data model mirrors the reference: cell-centred arrays carry ng ghost
 * layers and are filled by a periodic copy followed by physical boundary
 * conditions every RK stage; pencils are reconstructed into face states,
 * passed to the Riemann solver and differenced into `res`.
 *
 * Host-side exchange format (shared with the product's C-ABI): interior
 * cells only, structure-of-arrays, x fastest:
 *   prim[v][k][j][i]        v = 0..nv-1 (rho, v1, v2, [v3], p, scalars...)
 *   b_x1[k][j][i]  i in [0,nx1]   (x1-faces), b_x2[k][j][i] j in [0,nx2],
 *   b_x3[k][j][i]  k in [0,nx3]
 * and plane-parallel static profiles indexed by the x2 cell (or face) index.
 */
#ifndef PHLX_ORACLE_H
#define PHLX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LIM2ND = 0, ORC_PPH = 1, ORC_LIM5TH = 2 };
enum { ORC_HLLC = 0, ORC_LHLLC = 1, ORC_HLLD = 2, ORC_LHLLD = 3 };
enum { ORC_BC_PERIODIC = 0, ORC_BC_REFLECTIVE = 1, ORC_BC_OUTFLOW = 2, ORC_BC_DIODE = 3 };
enum { ORC_EOS_IDEAL = 0, ORC_EOS_HELMHOLTZ = 1, ORC_EOS_PIG = 2 };

typedef struct {
  int sdims, nx1, nx2, nx3, nas;
  double x1l, x1u, x2l, x2u, x3l, x3u;
  double gamma, mu;
  int recon, flux, all_mach, rk_stages;
  double cfl, mach_cutoff, mach_ct, tmax;
  int bc[6];              /* X1L X1U X2L X2U X3L X3U */
  int use_wb, use_gravity, evolve_etot, use_edot;
  int eos, fasteos;
  int nthreads;           /* OpenMP threads for the baseline (0 = default) */
  int shock_flattening;   /* USE_SHOCK_FLATTENING */
  double eps_sf;          /* eps_sf_make: relative pressure jump that flattens */
  int thermal_diffusion;  /* THERMAL_DIFFUSION_STS (RKL2, Strang-split around hydro_step) */
  int eval_parabolic_dt;  /* EVALUATE_PARABOLIC_TIMESTEP: dt_par every step (else at step 0) */
  double fix_t[6];        /* FIX_TEMPERATURE_AT_X1L..X3U wall temperature, NaN = not fixed */
} orc_config;

typedef struct orc_grid orc_grid;

orc_grid *orc_create(const orc_config *cfg);
void orc_destroy(orc_grid *g);
int orc_nvars(const orc_grid *g);

/* Interior SoA state in/out (see header comment for layout). Any pointer
 * may be NULL to skip that field. */
int orc_set_state(orc_grid *g, const double *prim, const double *eint,
                  const double *temp, const double *b_x1, const double *b_x2,
                  const double *b_x3);
int orc_get_state(const orc_grid *g, double *prim, double *eint, double *temp,
                  double *b_x1, double *b_x2, double *b_x3, double *b_cc);
/* Plane-parallel statics: cc arrays have nx2 entries, face arrays nx2+1. */
int orc_set_profiles(orc_grid *g, const double *eq_rho_cc, const double *eq_p_cc,
                     const double *eq_rho_f, const double *eq_p_f,
                     const double *grav_cc, const double *phi_cc,
                     const double *phi_f, const double *edot_cc);

/* FASTEOS well-balancing: equilibrium gamma_e at cell centres (nx2) and x2
 * faces (nx2+1). */
int orc_set_eos_profiles(orc_grid *g, const double *eq_gammae_cc, const double *eq_gammae_f);
/* Helmholtz EoS (eos = ORC_EOS_HELMHOLTZ, nas = 2 with Ye, 1/abar): the table
 * object is borrowed and must outlive the grid. */
struct orc_helm;
int orc_set_helm(orc_grid *g, const struct orc_helm *h);
/* gamma_e, gamma_c of the interior cells, out[2][cell] (x fastest). */
int orc_get_gammaf(const orc_grid *g, double *out);

/* time_loop step-0 initialisation: b_cc from faces, eint/T from (rho,p),
 * compute_hyperbolic_dt (:6012-6062, :6271-6409). */
int orc_init(orc_grid *g);
/* Resume after read_restart: eint (+ FASTEOS gammas) rebuilt from (rho, T)
 * or p, scalars from the file (:2761-2897, :6413-6500). */
int orc_resume(orc_grid *g, double time, double dt, double smax, long step);
/* USE_SHOCK_FLATTENING flags of the current state on cells -1..n per axis
 * (x3 only in 3D), x1 fastest; -1 when flattening is off. */
int orc_shock_flags(orc_grid *g, unsigned char *out);
/* Opacity kappa on cells -1..n of each axis (one ghost layer; x3 only in
 * 3D), x1 fastest, as the application fills lgrid%kappa. */
int orc_set_kappa(orc_grid *g, const double *kappa);
/* compute_parabolic_dt (:19804-19947); returns dt_par. */
double orc_parabolic_dt(orc_grid *g);
/* thermal_diffusion_step (:19949-20269) over half of the current dt. */
int orc_diffusion_step(orc_grid *g, int strang);
/* nsteps passes of the time_loop body: dt clip, [dt_par, diffusion 1],
 * hydro_step, [diffusion 2], time += dt, compute_hyperbolic_dt, step++
 * (:6662-6720). */
int orc_run(orc_grid *g, int nsteps);
/* One hydro_step with the current dt (no dt update). */
int orc_hydro_step(orc_grid *g);
int orc_compute_dt(orc_grid *g);
/* 0 time, 1 dt, 2 smax, 3 step */
double orc_get_scalar(const orc_grid *g, int which);
void orc_set_scalar(orc_grid *g, int which, double v);

/* Point kernels for unit tests: arrays are AoS [n][nv] / [n][sdims]. */
void orc_riemann_hydro(const orc_config *cfg, int n, const double *qL,
                       const double *qR, double *flux);
void orc_riemann_mhd(const orc_config *cfg, int n, const double *qL,
                     const double *qR, const double *bL, const double *bR,
                     double *flux, double *fluxb);
/* Reconstruct one pencil q[0..n-1] (n>=3): faces 1..n-2 of cells 1..n-2
 * (LIM5TH: cells 2..n-3); out qm[c] = left-edge value of cell c, qp[c] =
 * right-edge value. */
void orc_reconstruct(int scheme, int limited_scalar, int n, const double *q,
                     double *qm, double *qp);

/* ---- Helmholtz EoS (helm_oracle.c) ---- */
typedef struct orc_helm orc_helm;
/* f9_file / c4_file: the table in the reference file's column order,
 * [j (T)][i (rho)][9 | 4]. */
orc_helm *orc_helm_create(int nr, int nt, double ldlo, double ldhi, double ltlo,
                          double lthi, const double *f9_file, const double *c4_file);
void orc_helm_destroy(orc_helm *h);
double orc_helm_node(const orc_helm *h, int which, int idx);
/* PIG_EOS: the table is a partially-ionised-gas table (pig_rhoT_given_*,
 * source.F90:25457-26731): the same interpolation at rho (Ye = 1) with no
 * ion terms; abar and zbar are then ignored.  The Newton inversions are
 * the same (pig_tol = helm_tol). */
void orc_helm_set_pig(orc_helm *h, int pig);
void orc_helm_full(const orc_helm *h, double rho, double T, double abar, double zbar,
                   double *P, double *E, double *sound, double *cv);
void orc_helm_given1(const orc_helm *h, double rho, double T, double abar, double zbar,
                     double *P, double *dPdT, int ret);
void orc_helm_given2(const orc_helm *h, double rho, double T, double abar, double zbar,
                     double *E, double *dEdT, int ret);
int orc_helm_rhoe_given(const orc_helm *h, double rho, double E0, double abar, double zbar,
                        double *T, double *P, double *sound, int ret);
int orc_helm_rhoP_given(const orc_helm *h, double rho, double P0, double abar, double zbar,
                        double *T, double *E, double *sound, int ret);
void orc_helm_full_batch(const orc_helm *h, int n, const double *rho, const double *T,
                         const double *abar, const double *zbar, double *P, double *E,
                         double *cs, double *cv);
void orc_helm_rhoe_batch(const orc_helm *h, int n, const double *rho, const double *e,
                         const double *abar, const double *zbar, double *T, double *P,
                         double *cs, int ret, int *iters);
void orc_helm_rhoP_batch(const orc_helm *h, int n, const double *rho, const double *p,
                         const double *abar, const double *zbar, double *T, double *E,
                         double *cs, int ret, int *iters);
void orc_helm_interp(const orc_helm *h, double rhos, double T, double *out6);

#ifdef __cplusplus
}
#endif
#endif
