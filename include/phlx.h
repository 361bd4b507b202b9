/*
 * phlx.h -- C-ABI of the MI355X-native explicit finite-volume MHD update.
 *
 * The library replaces the device-side work of the reference's hydro_step
 * path behind plain pointers and sizes (no C++ or torch types), so a
 * Fortran caller binds it with ISO_C_BINDING and a Python caller with
 * ctypes.  Every entry point returns 0 on success or a negative PHLX_E_*
 * code; phlx_error_string() turns a code into text.  State lives on the GPU
 * in structure-of-arrays form; the library never keeps host pointers.
 * Calls on one context are ordered on that context's HIP stream; calls that
 * return host scalars synchronise it.  One context per host thread.
 *
 * Host exchange layout ("interior SoA"): interior cells only, x1 fastest:
 *   prim[v][k][j][i]   v = 0..nv-1 = rho, v1, v2, [v3], p, scalars
 *   b_x1[k][j][i] with i in [0,nx1], b_x2 with j in [0,nx2], b_x3 k in [0,nx3]
 * nv = sdims + 2 + nas (source/source.F90:472-477).
 * Static equilibrium / gravity / potential / heating inputs are
 * plane-parallel profiles indexed by the x2 cell (nx2 values) or x2 face
 * (nx2+1 values), the form every supported problem uses.
 */
#ifndef PHLX_H
#define PHLX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define PHLX_OK 0
#define PHLX_E_HIP (-1)        /* a HIP runtime call failed */
#define PHLX_E_CONFIG (-2)     /* unsupported or inconsistent configuration */
#define PHLX_E_ARG (-3)        /* null pointer / bad size / bad index */
#define PHLX_E_NOMEM (-4)      /* device allocation failed */
#define PHLX_E_STATE (-5)      /* call out of order (e.g. step before init) */
#define PHLX_E_NONFINITE (-6)  /* NaN / non-positive density or pressure found */
#define PHLX_E_TABLE (-7)      /* EoS table missing or malformed */

/* option values (names follow the reference's Makefile OPTS) */
enum { PHLX_LIM2ND_REC = 0, PHLX_PPH_REC = 1, PHLX_LIM5TH_REC = 2 };
enum { PHLX_HLLC_FLUX = 0, PHLX_LHLLC_FLUX = 1, PHLX_HLLD_FLUX = 2, PHLX_LHLLD_FLUX = 3 };
enum { PHLX_BC_PERIODIC = 0, PHLX_BC_REFLECTIVE = 1, PHLX_BC_OUTFLOW = 2, PHLX_BC_DIODE = 3 };
/* PHLX_EOS_PIG: PIG_EOS, a partially-ionised-gas table in the Helmholtz
 * table layout (pig_rhoT_given_*, source/source.F90:25457-26731), loaded with
 * phlx_load_pig_table; runs carry no advected scalars (nas 0) or two
 * (nas 2, as Helmholtz) that the EoS ignores. */
enum { PHLX_EOS_IDEAL = 0, PHLX_EOS_HELMHOLTZ = 1, PHLX_EOS_PIG = 2 };
enum { PHLX_PREC_FP64 = 0, PHLX_PREC_MIXED = 1 };

/* Run-time counterpart of the reference's compile-time OPTS subset
 * (Makefile.config:16-23, Make.checks:27-305, source.F90:31-477). */
typedef struct phlx_config {
  int32_t sdims;            /* sdims_make: 2 or 3 */
  int32_t nx1, nx2, nx3;    /* nx*_make (nx3 ignored for sdims=2) */
  int32_t nas;              /* nas_make: advected scalars */
  double x1l, x1u, x2l, x2u, x3l, x3u;  /* domain (initialize_simulation args) */
  double gamma, mu;         /* gamma_ad, mu */
  int32_t recon;            /* LIM2ND_REC / PPH_REC / LIM5TH_REC (ghosts are virtual: no ngc) */
  int32_t flux;             /* HLLC / LHLLC / HLLD / LHLLD (HLLD* imply USE_MHD) */
  int32_t all_mach;         /* FLUX_IS_ALL_MACH */
  int32_t rk_stages;        /* RK2_STEPPER (2) / RK3_STEPPER (3) */
  double cfl, mach_cutoff, mach_ct, tmax;
  int32_t bc[6];            /* X1L X1U X2L X2U X3L X3U */
  int32_t use_wb, use_gravity, evolve_etot, use_edot;
  int32_t eos, fasteos;     /* HELMHOLTZ_EOS / USE_FASTEOS */
  int32_t precision;        /* PHLX_PREC_* */
  int32_t device;           /* HIP device ordinal */
  int32_t reserved[8];      /* must be zero */
} phlx_config;

typedef struct phlx_ctx phlx_ctx;

const char *phlx_error_string(int code);
int phlx_version(void);

/* Allocation part of initialize_simulation (source/source.F90:1150-1881):
 * validates cfg, allocates every device buffer, creates the context's
 * stream. */
int phlx_create(const phlx_config *cfg, phlx_ctx **out);
/* finalize_simulation (source/source.F90:1889): frees device memory. */
int phlx_destroy(phlx_ctx *ctx);
int phlx_nvars(const phlx_ctx *ctx);

/* Upload / download of the state held in the reference's `locgrid`
 * (prim, eint, temp, b_x1..3; source/source.F90:839-1130) in interior-SoA
 * layout.  Any pointer may be NULL to skip that field.  b_cc (download
 * only) is 3 x N. */
int phlx_upload_state(phlx_ctx *ctx, const double *prim, const double *eint,
                      const double *temp, const double *b_x1, const double *b_x2,
                      const double *b_x3);
int phlx_download_state(phlx_ctx *ctx, double *prim, double *eint, double *temp,
                        double *b_x1, double *b_x2, double *b_x3, double *b_cc);
/* The reference's own array layout: prim(1:nv, lx1-ngc:ux1+ngc,
 * lx2-ngc:ux2+ngc, lx3-ngc:ux3+ngc) in Fortran order (variable fastest,
 * ghosts included; in 2D the x3 extent is 1).  Ghost values are ignored on
 * upload and written as zeros on download. */
int phlx_upload_prim_aos(phlx_ctx *ctx, const double *prim, int32_t ngc);
int phlx_download_prim_aos(phlx_ctx *ctx, double *prim, int32_t ngc);
/* Plane-parallel statics: eq_prim_cc / eq_prim_x2 (rho, p), grav(2,:),
 * phi_cc / phi_x2, edot (source/source.F90:1415-1483). NULL = zeros. */
int phlx_set_profiles(phlx_ctx *ctx, const double *eq_rho_cc, const double *eq_p_cc,
                      const double *eq_rho_f, const double *eq_p_f,
                      const double *grav_cc, const double *phi_cc,
                      const double *phi_f, const double *edot_cc);

/* USE_FASTEOS + USE_WB: equilibrium gamma_e = p_eq/e_eq + 1 at x2 cell
 * centres (nx2) and x2 faces (nx2+1); the reference's eq_gammae_cc / _x1 /
 * _x2 / _x3 (source/source.F90:1465-1478) reduced to plane-parallel
 * profiles (x1 and x3 faces sit at cell-centre heights). */
int phlx_set_eos_profiles(phlx_ctx *ctx, const double *eq_gammae_cc,
                          const double *eq_gammae_f);
/* gamma_e, gamma_c of the interior cells (the reference's gammaf(1:2,...)),
 * host array [2][nx3][nx2][nx1]; Helmholtz contexts only. */
int phlx_download_gammaf(phlx_ctx *ctx, double *gammaf);

/* Step-0 part of time_loop (source/source.F90:6012-6062, 6271-6409):
 * b_cc from the face field, eint and T from (rho, p), then
 * compute_hyperbolic_dt.  Sets time = 0, step = 0. */
int phlx_init(phlx_ctx *ctx);
/* Resume from a checkpoint, the read_restart path (source/source.F90
 * :2761-2897) plus time_loop's step /= 0 branch (:6413-6500): after
 * phlx_upload_state of the restart file's prim, T and face fields, rebuilds
 * b_cc, eint (Helmholtz: a full evaluation at (rho, T); ideal: p/(gamma-1))
 * and the FASTEOS gammas, and takes time, dt, smax and step as given (dt is
 * not recomputed).  dt, smax > 0, step >= 0. */
int phlx_resume(phlx_ctx *ctx, double time, double dt, double smax, int64_t step);
/* hydro_step(mgrid, lgrid) (source/source.F90:6847-13701) with the current
 * dt; ghost zones are produced on the device (no halo exchange). */
int phlx_hydro_step(phlx_ctx *ctx);
/* compute_hyperbolic_dt (source/source.F90:13707-13960): smax = max over
 * zones of sum_d (|v_d| + c_f)/dx_d, dt = cfl/smax.  Outputs may be NULL
 * (then the call does not synchronise). */
int phlx_compute_dt(phlx_ctx *ctx, double *dt_out, double *smax_out);
/* nsteps passes of the time_loop body (source/source.F90:6662-6720): clip
 * dt to tmax, hydro_step, time += dt, compute_hyperbolic_dt, step += 1.
 * Entirely device-resident and asynchronous: no host read-back per step. */
int phlx_run(phlx_ctx *ctx, int32_t nsteps);
int phlx_synchronize(phlx_ctx *ctx);
int phlx_get_scalars(phlx_ctx *ctx, double *time, double *dt, double *smax,
                     int64_t *step);
int phlx_set_scalars(phlx_ctx *ctx, double time, double dt, double smax, int64_t step);

/* Fine-grained form of hydro_step for callers that keep the RK stage loop
 * (source/source.F90:7149-13699): stage_begin(irk) rebuilds state0 from
 * prim/eint at irk == 0 (:7057-7147); sweep(dir) does the directional
 * sweep dir = 0,1,2 (x1 :7960-9105, x2 :9107-10301, x3 :10303-11364);
 * ct() the corner EMFs and face-B update (:11966-12437); rk_update() the
 * sources, RK combine and cons->prim (:12606-13686).  (a1, a2, a3) is the
 * stage's Shu-Osher row; a3 is multiplied by the device-resident dt.
 * With RK2 and MHD the two ct() calls of a step come in stage order (the
 * stage-1 face field is held in the second face buffer between them);
 * ct(1) before ct(0), or ct(0) twice, returns PHLX_E_STATE. */
int phlx_stage_begin(phlx_ctx *ctx, int32_t irk);
int phlx_sweep(phlx_ctx *ctx, int32_t dir);
int phlx_ct(phlx_ctx *ctx, int32_t irk, double a1, double a2, double a3);
int phlx_rk_update(phlx_ctx *ctx, int32_t irk, double a1, double a2, double a3,
                   int32_t last_stage);
/* Whether phlx_hydro_step / phlx_run fuse each stage's three directional
 * sweeps and flux divergence (the work of sweep(0..2) and the divergence
 * and RK combine of rk_update) into one pass that stages a pencil tile in
 * LDS and keeps the face fluxes out of HBM: mode 1 wherever the run admits
 * it, 0 never, -1 (the default) where it is the faster form: runs without
 * MHD.  It applies to 3D runs with LIM2ND or PPH, the ideal or
 * the FASTEOS Helmholtz state, no shock flattening and nx1, nx2 multiples
 * of 16 (other runs take the separate sweeps).  Either way the results are
 * bit for bit the same.  No counterpart in the reference (its hydro_step is
 * one routine either way): for A/B timing and for callers that want the
 * per-direction kernels. */
int phlx_set_fused_sweep(phlx_ctx *ctx, int32_t mode);
/* Whether phlx_run replays one captured HIP graph per step (on = 1) instead
 * of launching the step's kernels one by one (0, the default).  The step is
 * captured on the first phlx_run that needs it and again whenever a setting
 * that changes its kernels or their arguments has changed; runs with
 * thermal diffusion or profiling launch eagerly.  The results are bit for
 * bit the same.  Measured on MI355X it is not faster (2D 256^2, where
 * each step is a dozen short kernels: 0.090 against 0.085 ms per step), so
 * it is off by default; it is for callers that embed the step in their own
 * graphs or want one launch per step.  No counterpart in the reference. */
int phlx_set_step_graph(phlx_ctx *ctx, int32_t on);

/* ---- Helmholtz EoS ----
 * load_helm_table replaces load_helm_table (source/source.F90:23077-23156):
 * f9 / c4 are the table's f block and dP/d(rho Ye) block in the reference
 * file's column order (f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt and
 * c, cd, ct, cdt), host memory laid out like the reference's
 * helm_table_var(9, nrho, nT) in Fortran order (column fastest, then rho,
 * then T).  Nodes are rho_i = 10**(ldlo + i*(ldhi-ldlo)/(nrho-1)) and
 * T_j = 10**(ltlo + j*(lthi-ltlo)/(nT-1)) (:23094-23104).  The library keeps
 * its own device copy; the host arrays may be freed on return. */
int phlx_load_helm_table(phlx_ctx *ctx, const double *f9, const double *c4, int32_t nrho,
                         int32_t nT, double ldlo, double ldhi, double ltlo, double lthi);
/* load_pig_table (source/source.F90:25407-25455): a PIG_EOS table, same
 * arguments and column order as phlx_load_helm_table; contexts of eos
 * PHLX_EOS_PIG take only this one (and Helmholtz contexts only the other),
 * a batch context of the ideal gas either. */
int phlx_load_pig_table(phlx_ctx *ctx, const double *f9, const double *c4, int32_t nrho,
                         int32_t nT, double ldlo, double ldhi, double ltlo, double lthi);
/* Batched helm_rhoT_given_full (source/source.F90:23686-24232): n points,
 * all pointers DEVICE memory of the context's GPU; P, E (specific internal
 * energy, erg/g), sound speed (relativistic form) and c_v; any output may
 * be NULL.  Asynchronous on the context stream. */
int phlx_helm_rhoT_full_batch(phlx_ctx *ctx, int64_t n, const double *rho, const double *T,
                              const double *abar, const double *zbar, double *P, double *E,
                              double *sound, double *cv);
/* Batched helm_rhoe_given (source/source.F90:24982-25019): Newton for
 * T(rho, e) from the warm start in T (in/out), <= 100 iterations, tol 1e-11,
 * T outside [10**ltlo, 10**lthi] resets to the start value.  Returns P and,
 * when return_sound != 0, the sound speed (else sound = 0).  iters (may be
 * NULL) receives the iterations taken per point.  Device pointers. */
int phlx_helm_rhoe_given_batch(phlx_ctx *ctx, int64_t n, const double *rho, const double *e,
                               const double *abar, const double *zbar, double *T, double *P,
                               double *sound, int32_t return_sound, int32_t *iters);
/* Batched helm_rhoP_given (source/source.F90:24943-24980): as above for
 * T(rho, P); returns E. */
int phlx_helm_rhoP_given_batch(phlx_ctx *ctx, int64_t n, const double *rho, const double *p,
                               const double *abar, const double *zbar, double *T, double *E,
                               double *sound, int32_t return_sound, int32_t *iters);

/* ---- T(rho, e) surrogate: guess of the Newton inversion ----
 * No counterpart in the reference, whose helm_rhoe_given
 * (source/source.F90:24982-25019) starts from the carried temperature.  A
 * 4 -> 64 -> 64 -> 1 tanh network maps normalised (log10 rho, log10 e,
 * abar, zbar) to normalised log10 T; the 64 x 64 layer runs on BF16 MFMA
 * with FP32 accumulation.  Parameters are a flat FP32 vector of 4545 values
 * (w1[4][64], b1[64], w2[64][64], b2[64], w3[64], b3) and a normalisation
 * norm[10] = (in_off[4], in_scale[4], out_off, out_scale):
 * x = (raw - in_off) * in_scale, log10 T = y * out_scale + out_off.
 *
 * phlx_mlp_train: generates nsamples (rho, e, abar, zbar) -> T pairs with
 * the loaded table (log-uniform over the training box, by default the table
 * box with abar in [4, 16] and Ye in [0.45, 0.5]; labels from the FP64
 * Newton solve) and trains on the GPU
 * (Adam, cosine learning rate, loss weighted by (d ln E / d ln T)^2, i.e.
 * measured in the energy residual).  Outputs (may be NULL): the last batch
 * loss and the held-out RMS errors of log10 T and of the implied log10 e. */
/* Training box of later phlx_mlp_train calls, box[8] = {log10(rho*Ye) lo, hi,
 * log10 T lo, hi, abar lo, hi, Ye lo, hi} (host array; clipped to the table;
 * NULL restores the table box).  A network trained on the states a run
 * occupies is far more accurate there than one spread over the table's 27
 * decades of density. */
int phlx_mlp_set_train_box(phlx_ctx *ctx, const double *box);
int phlx_mlp_train(phlx_ctx *ctx, int64_t nsamples, int32_t epochs, int32_t batch, double lr,
                   uint64_t seed, double *train_loss, double *val_rms_logT,
                   double *val_rms_loge);
int phlx_mlp_set_params(phlx_ctx *ctx, const float *theta, const float *norm);
int phlx_mlp_get_params(phlx_ctx *ctx, float *theta, float *norm);
/* Inference on raw inputs raw[n][4] (device): normalised outputs y[n] and
 * in-box flags (may be NULL). */
int phlx_mlp_forward(phlx_ctx *ctx, int64_t n, const float *raw, float *y, int32_t *inbox);
/* Gradient of (1/n) sum wt (y - tn)^2 over normalised inputs xn[n][4] (all
 * device; wt may be NULL for unit weights) into grad[4545] (device); loss to
 * the host.  For checking the training kernels. */
int phlx_mlp_grad(phlx_ctx *ctx, int64_t n, const float *xn, const float *tn, const float *wt,
                  float *grad, double *loss);
/* helm_rhoe_given with a chosen start: guess_mode 0 = the given T (the
 * reference), 1 = the surrogate (falls back to the bracketed guess when the
 * inputs leave its training box or its first energy residual exceeds the
 * threshold of phlx_set_eos_guess, default 0.25), 2 = bracketed guess
 * (bisection in log T over the table).  fallback (may be NULL) flags points
 * that fell back.  Device pointers. */
int phlx_helm_rhoe_given_guess_batch(phlx_ctx *ctx, int64_t n, const double *rho,
                                     const double *e, const double *abar, const double *zbar,
                                     double *T, double *P, double *sound, int32_t return_sound,
                                     int32_t guess_mode, int32_t *iters, int32_t *fallback);
/* Start of the per-zone T(rho, e) Newton inside the Helmholtz update: 0 the
 * carried T (the reference), 1 the carried T corrected by the surrogate
 * (log10 T_prev + s (y_now - y_prev)), 2 bracketed, 3 the energy predictor
 * T_prev + (e - e_prev) / c_v of the zone's previous solve (its target e and
 * c_v kept per zone; a new state, a diffusion step or a mode change start
 * each zone from the carried T once).  The Newton then converges to the
 * reference's tolerance from that start.  count_iterations != 0
 * accumulates Newton iterations / solves / fallbacks (phlx_eos_stats).
 * Modes 1 and 2, and counting, need nx1 % 64 == 0 and nx2 % 4 == 0.
 * resid_tol is also the surrogate threshold of
 * phlx_helm_rhoe_given_guess_batch; a context of another EoS may set it once
 * a Helmholtz table is loaded. */
int phlx_set_eos_guess(phlx_ctx *ctx, int32_t mode, double resid_tol, int32_t count_iterations);
/* Mixed precision with the Helmholtz EoS: by default (lin_tol 0) the
 * per-zone T(rho, e) solve iterates the reference's Newton to its tolerance
 * (helm_rhoe_given, source/source.F90:24982-25019), one FP64 evaluation per
 * iterate.  lin_tol > 0 selects a shortcut with no counterpart in the
 * reference: one FP64 evaluation at the start and, when |e(T) - e| / e <
 * lin_tol, the Newton step applied linearly to T and P instead of iterating
 * (error second order in the residual).  lin_tol in [0, 1e-2]. */
int phlx_set_eos_lin_tol(phlx_ctx *ctx, double lin_tol);
int phlx_eos_stats(phlx_ctx *ctx, int64_t *newton_iters, int64_t *solves, int64_t *fallbacks,
                   int32_t reset);
/* THERMAL_DIFFUSION_STS (compile-time OPTS in the reference, with
 * EVALUATE_PARABOLIC_TIMESTEP and FIX_TEMPERATURE_AT_X1L..X3U; time_loop
 * source/source.F90:6672-6688, thermal_diffusion_step :19949-20269).
 * Radiative diffusion of the internal energy, flux (4/3) a c K_int dT/dx
 * with K = T^3 / (rho kappa) (harmonic mean on faces), by RKL2
 * super-time-stepping: phlx_run then runs diffusion over dt/2, hydro_step,
 * diffusion over dt/2, and the hyperbolic dt of the result.  evaluate_dt 1:
 * dt_par = 1 / max(2 c / (3 kappa rho) sum 1/dx^2) every step, 0: once in
 * phlx_init (set kappa first).  fix_t: NULL or six wall temperatures
 * X1L X1U X2L X2U X3L X3U, NaN for the zero-gradient ghost; a fixed one
 * sets the ghost to 2 T_wall - T (non-periodic sides only).  The ideal gas
 * needs nas 0 (PHLX_E_CONFIG otherwise); diffusion evolves eint (no
 * STS_EVOLVE_TEMP), Cartesian, no USE_TIMMES_KAPPA, no internal boundaries. */
int phlx_set_thermal_diffusion(phlx_ctx *ctx, int32_t enable, int32_t evaluate_dt,
                               const double *fix_t);
/* Opacity on cells -1..n of each axis (one ghost layer; the x3 layer only in
 * 3D), x1 fastest: (nx1+2)*(nx2+2)*(nx3+2 or 1) positive doubles, as the
 * application fills lgrid%kappa (source/source.F90:1523).  After
 * phlx_set_thermal_diffusion. */
int phlx_set_kappa(phlx_ctx *ctx, const double *kappa);
/* compute_parabolic_dt (source/source.F90:19804-19947); dt_par may be NULL. */
int phlx_parabolic_dt(phlx_ctx *ctx, double *dt_par);
/* thermal_diffusion_step(strang) over half the current dt (after
 * phlx_init), for callers that drive the step themselves. */
int phlx_diffusion_step(phlx_ctx *ctx, int32_t strang);
/* RKL2 stage count of the last diffusion step (the reference prints it). */
int phlx_diffusion_stages(phlx_ctx *ctx, int32_t *stages);
/* USE_SHOCK_FLATTENING (a compile-time OPTS switch with eps_sf_make in the
 * reference; source/source.F90:8073-8131 detection, :8731-8740 and
 * :8743-8869 use).  At the start of every stage a cell whose summed relative
 * pressure jump sum_d |p(c+e_d) - p(c-e_d)| / p(c) exceeds eps_sf (the
 * reference's default 0.25) flags its 3^sdims neighbourhood; a flagged cell
 * is reconstructed with the LIM2ND slope, and a face beside one takes the
 * two-wave HLL flux (HLLC :15929-16003, HLLD :14386-14486).  enable 0 turns
 * it off.  FP64 contexts only (PHLX_E_CONFIG otherwise), and every axis
 * holds at least 3 cells.  May be called before or after phlx_init. */
int phlx_set_shock_flattening(phlx_ctx *ctx, int32_t enable, double eps_sf);

/* Diagnostics: max over cells of |sum_d (b_d(+) - b_d(-))/dx_d|. */
int phlx_divb_max(phlx_ctx *ctx, double *out);

/* phlx_compute_dt without a context and without allocating: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc).  The same
 * signal-speed maximum as phlx_compute_dt (compute_hyperbolic_dt,
 * source/source.F90:13707-13960), ideal-gas EoS only, over caller-owned
 * device arrays: prim [nv][nx3][nx2][nx1] (interior SoA), bcc [sdims][...]
 * (cell-centred B; MHD only, else may be NULL); writes dt = cfl / smax and
 * smax to dt_smax[0..1] (device memory), asynchronously on `stream`
 * (a hipStream_t; NULL = the null stream).  Partial maxima go to a
 * library-owned device array of 2048 doubles per GPU (16 KiB, held until the
 * process ends), so any grid up to 2^33 cells fits; larger grids are
 * rejected (PHLX_E_ARG).  Calls that overlap on one GPU race on that array:
 * serialise them (one stream, or events).  cfg must describe the arrays
 * (sdims, nx*, domain, gamma, cfl, flux for MHD, eos = PHLX_EOS_IDEAL). */
int phlx_compute_dt_noalloc(const phlx_config *cfg, const double *prim, const double *bcc,
                            double *dt_smax, void *stream);
/* Per-kernel device time accumulated while profiling is on (HIP events on
 * the context stream).  kernel ids: 0 flux x1, 1 flux x2, 2 flux x3,
 * 3 corner EMF, 4 face-B, 5 update, 6 misc (shock flags), 7 thermal
 * diffusion. */
int phlx_profile_enable(phlx_ctx *ctx, int32_t on);
int phlx_profile_read(phlx_ctx *ctx, int32_t kernel, double *total_ms, int64_t *launches);
/* Bandwidth probe: copies nbytes between two fresh device buffers `reps`
 * times with a plain 16-byte-vector copy kernel and returns the achieved
 * HBM rate (read + write bytes per second, in GB/s).  Diagnostic only. */
int phlx_stream_copy(phlx_ctx *ctx, int64_t nbytes, int32_t reps, double *gbps);
/* Opaque HIP stream of the context (hipStream_t), for callers that want
 * to order their own work after the library's. */
void *phlx_stream(phlx_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
