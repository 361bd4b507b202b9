// phlx_api.hip -- the C-ABI (include/phlx.h): context, device buffers,
// host exchange, stepping and profiling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/phlx.h"
#include "phlx_internal.hpp"

using namespace phlx;

struct ProfRec {
  int kid;
  hipEvent_t a, b;
};

struct phlx_ctx {
  phlx_config cfg;
  Grid g;
  KernelTable kt;
  hipStream_t stream = nullptr;
  int device = 0;
  int stages = 2;
  double rk[3][3];
  bool initialized = false;
  std::vector<void *> allocs;
  double *scratch_val = nullptr;
  // Helmholtz table (device records + node arrays), owned separately so a
  // table can be replaced
  void *helm_buf = nullptr;
  HelmDev helm{};
  bool helm_loaded = false;
  // T(rho, e) surrogate: parameters, Adam moments, BF16 copies, gradient
  // partials (one allocation, made on first use)
  void *mlp_buf = nullptr;
  float *mlp_theta = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_w1q = nullptr,
        *mlp_w3q = nullptr, *mlp_gpart = nullptr, *mlp_lpart = nullptr, *mlp_gsum = nullptr;
  unsigned short *mlp_wa = nullptr;
  bool mlp_ready = false;
  // training box {log10 rhoYe lo, hi, log10 T lo, hi, abar lo, hi, Ye lo, hi};
  // has_box false: the table's range
  double mlp_box[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool mlp_has_box = false;
  bool eos_stats = false;
  // THERMAL_DIFFUSION_STS: on / dt_par every step (else at step 0), the
  // host copy of dt_par, kappa uploaded, the ideal gas' diffusion T current
  // (false after any hydro update: the next step recomputes it from p)
  bool diff_on = false, diff_eval = true, kappa_set = false, tdiff_valid = false;
  // mixed path: a step was launched since the last rebuild of the FP64
  // prim / eint / b_cc from U (whether it ran, and so whether they lag U,
  // the device flags SC_U_CUR / SC_P_STALE say)
  bool prim_stale = false;
  // hydro_step's fused sweep: -1 = where it is the faster form, 0 = never,
  // 1 = wherever the kernel family has one
  int fused_sweep = -1;
  // RK2 with MHD: between a step's two CT calls the stage-1 face field
  // lives in the b0f buffers and g.bf / g.b0f are swapped (k_faceb)
  bool b_swapped = false;
  // phlx_run's captured step (phlx_set_step_graph) and what it was
  // captured with: the grid descriptor (every kernel argument but the
  // per-stage constants), the sweep form and the Newton counting
  bool step_graph = false;
  hipGraph_t sg_graph = nullptr;
  hipGraphExec_t sg_exec = nullptr;
  Grid sg_g{};
  bool sg_fused = false, sg_stats = false;
  double dt_par = 0.0;
  int diff_stages = 0;  // stages of the last diffusion step
  // profiling
  bool prof = false;
  std::vector<ProfRec> recs;
  double prof_ms[KID_COUNT] = {0};
  long long prof_n[KID_COUNT] = {0};
};

#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "phlx: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                        \
      return PHLX_E_HIP;                                                       \
    }                                                                          \
  } while (0)

extern "C" const char *phlx_error_string(int code) {
  switch (code) {
    case PHLX_OK: return "ok";
    case PHLX_E_HIP: return "HIP runtime error";
    case PHLX_E_CONFIG: return "unsupported or inconsistent configuration";
    case PHLX_E_ARG: return "invalid argument";
    case PHLX_E_NOMEM: return "device allocation failed";
    case PHLX_E_STATE: return "call out of order";
    case PHLX_E_NONFINITE: return "non-finite or non-positive state";
    case PHLX_E_TABLE: return "EoS table missing or malformed";
  }
  return "unknown error";
}

extern "C" int phlx_version(void) { return 100; }

// Validation of the supported option subset, mirroring the reference's
// Make.checks rules that apply (exactly one stepper/recon/flux, HLLD
// requires MHD, periodic in pairs, FASTEOS requires a complex EoS).
// plane-parallel profile slots: eq_rho_cc, eq_p_cc, eq_rho_f, eq_p_f, grav,
// phi_cc, phi_f, edot, eq_gammae_cc, eq_gammae_f; each ny+1 doubles
constexpr int kNumProfiles = 10;

static int validate(const phlx_config *c) {
  if (c->sdims != 2 && c->sdims != 3) return PHLX_E_CONFIG;
  if (c->nx1 < 2 || c->nx2 < 2 || (c->sdims == 3 && c->nx3 < 2)) return PHLX_E_CONFIG;
  if (c->recon != PHLX_LIM2ND_REC && c->recon != PHLX_PPH_REC && c->recon != PHLX_LIM5TH_REC)
    return PHLX_E_CONFIG;
  if (c->flux < PHLX_HLLC_FLUX || c->flux > PHLX_LHLLD_FLUX) return PHLX_E_CONFIG;
  if (c->rk_stages != 2 && c->rk_stages != 3) return PHLX_E_CONFIG;
  if (c->nas != 0 && c->nas != 2) return PHLX_E_CONFIG;
  const bool mhd = c->flux >= PHLX_HLLD_FLUX;
  for (int d = 0; d < c->sdims; d++) {
    const int lo = c->bc[2 * d], hi = c->bc[2 * d + 1];
    if (lo < 0 || lo > 3 || hi < 0 || hi > 3) return PHLX_E_CONFIG;
    if ((lo == PHLX_BC_PERIODIC) != (hi == PHLX_BC_PERIODIC)) return PHLX_E_CONFIG;
    if (mhd && (lo == PHLX_BC_DIODE || hi == PHLX_BC_DIODE)) return PHLX_E_CONFIG;
  }
  // Helmholtz runs carry Ye and 1/abar (ADVECT_YE_IABAR) as the two
  // advected scalars; without USE_FASTEOS the Riemann solvers make a
  // T(rho, P) Newton per face state (FP64 only); FASTEOS needs a complex EoS
  // (Make.checks)
  // PIG_EOS runs on the same tabulated-EoS kernels (its table has Ye = 1
  // and no ion terms, so the two advected scalars are carried unused)
  if (c->eos == PHLX_EOS_HELMHOLTZ || c->eos == PHLX_EOS_PIG) {
    if (c->nas != 2 && !(c->eos == PHLX_EOS_PIG && c->nas == 0)) return PHLX_E_CONFIG;
    if (!c->fasteos && c->precision != PHLX_PREC_FP64) return PHLX_E_CONFIG;
  } else if (c->eos == PHLX_EOS_IDEAL) {
    if (c->fasteos) return PHLX_E_CONFIG;
  } else {
    return PHLX_E_CONFIG;
  }
  if (c->precision != PHLX_PREC_FP64 && c->precision != PHLX_PREC_MIXED) return PHLX_E_CONFIG;
  if (c->cfl <= 0.0 || c->gamma <= 1.0 || c->mu <= 0.0) return PHLX_E_CONFIG;
  for (int r = 0; r < 8; r++)
    if (c->reserved[r] != 0) return PHLX_E_CONFIG;
  return PHLX_OK;
}

template <typename T>
static int dalloc(phlx_ctx *ctx, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, n * sizeof(T) + 256) != hipSuccess) return PHLX_E_NOMEM;
  if (hipMemsetAsync(q, 0, n * sizeof(T) + 256, ctx->stream) != hipSuccess) return PHLX_E_HIP;
  ctx->allocs.push_back(q);
  *p = (T *)q;
  return PHLX_OK;
}

extern "C" int phlx_create(const phlx_config *cfg, phlx_ctx **out) {
  if (!cfg || !out) return PHLX_E_ARG;
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  phlx_ctx *ctx = new (std::nothrow) phlx_ctx();
  if (!ctx) return PHLX_E_NOMEM;
  ctx->cfg = *cfg;
  ctx->device = cfg->device;
  if (hipSetDevice(cfg->device) != hipSuccess) {
    delete ctx;
    return PHLX_E_HIP;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return PHLX_E_HIP;
  }
  Grid &g = ctx->g;
  std::memset(&g, 0, sizeof(Grid));
  g.sd = cfg->sdims;
  g.nas = cfg->nas;
  g.nv = cfg->sdims + 2 + cfg->nas;
  g.ip = cfg->sdims + 1;
  g.n1 = cfg->nx1;
  g.n2 = cfg->nx2;
  g.n3 = cfg->sdims == 3 ? cfg->nx3 : 1;
  g.N = (long long)g.n1 * g.n2 * g.n3;
  g.nf[0] = (long long)(g.n1 + 1) * g.n2 * g.n3;
  g.nf[1] = (long long)g.n1 * (g.n2 + 1) * g.n3;
  g.nf[2] = (long long)g.n1 * g.n2 * (g.n3 + 1);
  for (int i = 0; i < 6; i++) g.bc[i] = cfg->bc[i];
  // source/source.F90:1725-1734
  const double dx1 = (cfg->x1u - cfg->x1l) / (double)cfg->nx1;
  const double dx2 = (cfg->x2u - cfg->x2l) / (double)cfg->nx2;
  const double dx3 = cfg->sdims == 3 ? (cfg->x3u - cfg->x3l) / (double)cfg->nx3 : 1.0;
  g.inv_dx[0] = 1.0 / dx1;
  g.inv_dx[1] = 1.0 / dx2;
  g.inv_dx[2] = 1.0 / dx3;
  g.gm = cfg->gamma;
  g.gmm1 = g.gm - 1.0;
  g.igmm1 = 1.0 / g.gmm1;
  g.inv_mu = 1.0 / cfg->mu;
  g.mach_cutoff = cfg->mach_cutoff;
  g.mach_ct = cfg->mach_ct;
  g.recon = cfg->recon;
  g.flux = cfg->flux;
  g.all_mach = cfg->all_mach;
  g.wb = cfg->use_wb;
  g.grav = cfg->use_gravity;
  g.etot = cfg->evolve_etot;
  g.edot = cfg->use_edot;
  g.fasteos = cfg->fasteos;
  g.mhd = cfg->flux >= PHLX_HLLD_FLUX;
  ctx->stages = cfg->rk_stages;
  // Shu-Osher rows (source/source.F90:7018-7048)
  if (ctx->stages == 2) {
    const double r[2][3] = {{1.0, 0.0, -1.0}, {0.5, 0.5, -0.5}};
    for (int s = 0; s < 2; s++)
      for (int c = 0; c < 3; c++) ctx->rk[s][c] = r[s][c];
  } else {
    ctx->rk[0][0] = 1.0; ctx->rk[0][1] = 0.0; ctx->rk[0][2] = -1.0;
    ctx->rk[1][0] = 0.75; ctx->rk[1][1] = 0.25; ctx->rk[1][2] = -0.25;
    ctx->rk[2][0] = 1.0 / 3.0; ctx->rk[2][1] = 2.0 / 3.0; ctx->rk[2][2] = -(2.0 / 3.0);
  }
  const bool helm = cfg->eos == PHLX_EOS_HELMHOLTZ || cfg->eos == PHLX_EOS_PIG;
  if (!select_kernels(g.sd, g.nas, g.mhd, helm, cfg->fasteos != 0,
                      cfg->precision == PHLX_PREC_MIXED, cfg->recon == PHLX_LIM5TH_REC,
                      ctx->kt)) {
    phlx_destroy(ctx);
    return PHLX_E_CONFIG;
  }
  const size_t N = (size_t)g.N, nv = (size_t)g.nv;
  int e = 0;
  e |= dalloc(ctx, &g.P, nv * N);
  e |= dalloc(ctx, &g.U, nv * N);
  e |= dalloc(ctx, &g.U0, nv * N);
  e |= dalloc(ctx, &g.eint, N);
  if (cfg->precision == PHLX_PREC_MIXED) e |= dalloc(ctx, &g.Pf, (nv + 5) * N);
  for (int d = 0; d < g.sd; d++) {
    e |= dalloc(ctx, &g.F[d], nv * (size_t)g.nf[d]);
    g.Ff[d] = (float *)g.F[d];  // the mixed path stores FP32 fluxes in the same buffer
  }
  if (g.mhd) {
    e |= dalloc(ctx, &g.bcc, 3 * N);
    e |= dalloc(ctx, &g.ecell, 3 * N);
    // the mixed path stores the cell and face EMFs as FP32 in the same buffers
    g.ecellf = cfg->precision == PHLX_PREC_MIXED ? (float *)g.ecell : nullptr;
    for (int d = 0; d < g.sd; d++) {
      e |= dalloc(ctx, &g.bf[d], (size_t)g.nf[d]);
      e |= dalloc(ctx, &g.b0f[d], (size_t)g.nf[d]);
      e |= dalloc(ctx, &g.emf[d], 2 * (size_t)g.nf[d]);
      g.emff[d] = (float *)g.emf[d];
      e |= dalloc(ctx, &g.sgn[d], (size_t)g.nf[d]);
    }
    const size_t n1 = g.n1, n2 = g.n2, n3 = g.n3;
    e |= dalloc(ctx, &g.cor[2], (n1 + 1) * (n2 + 1) * n3);
    if (g.sd == 3) {
      e |= dalloc(ctx, &g.cor[0], n1 * (n2 + 1) * (n3 + 1));
      e |= dalloc(ctx, &g.cor[1], (n1 + 1) * n2 * (n3 + 1));
    }
    for (int d = 0; d < 3; d++) g.corf[d] = (float *)g.cor[d];   // mixed: FP32 edge EMFs
  }
  if (helm) {
    e |= dalloc(ctx, &g.temp, N);
    e |= dalloc(ctx, &g.gam, 2 * N);
  }
  g.eos_guess = GUESS_WARM;
  // the mixed path converges the T(rho, e) Newton to the reference's
  // tolerance; phlx_set_eos_lin_tol(> 0) selects the linearised shortcut
  g.eos_lin_tol = 0.0;
  g.mlp.resid_tol = 0.25f;
  g.mlp.box = 1.02f;
  double *prof = nullptr;
  const size_t ny = g.n2;
  e |= dalloc(ctx, &prof, kNumProfiles * (ny + 1));
  e |= dalloc(ctx, &g.sc, 1);
  e |= dalloc(ctx, &ctx->scratch_val, 1);
  g.nred = cell_grid_blocks(g);
  e |= dalloc(ctx, &g.red, (size_t)g.nred);
  if (helm) {
    e |= dalloc(ctx, &g.eos_part, 3 * (size_t)g.nred);
    e |= dalloc(ctx, &g.mlp_y, N);
    e |= dalloc(ctx, &g.eprev, N);
    e |= dalloc(ctx, &g.cvf, N);
  }
  if (e) {
    phlx_destroy(ctx);
    return PHLX_E_NOMEM;
  }
  g.eq_rho_cc = prof + 0 * (ny + 1);
  g.eq_p_cc = prof + 1 * (ny + 1);
  g.eq_rho_f = prof + 2 * (ny + 1);
  g.eq_p_f = prof + 3 * (ny + 1);
  g.gravy = prof + 4 * (ny + 1);
  g.phi_cc = prof + 5 * (ny + 1);
  g.phi_f = prof + 6 * (ny + 1);
  g.edot_cc = prof + 7 * (ny + 1);
  g.eq_ge_cc = prof + 8 * (ny + 1);
  g.eq_ge_f = prof + 9 * (ny + 1);
  Scalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.tmax = cfg->tmax;
  s0.cfl = cfg->cfl;
  s0.active = 1;
  if (hipMemcpyAsync(g.sc, &s0, sizeof(s0), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    phlx_destroy(ctx);
    return PHLX_E_HIP;
  }
  *out = ctx;
  return PHLX_OK;
}

extern "C" int phlx_destroy(phlx_ctx *ctx) {
  if (!ctx) return PHLX_E_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->sg_exec) (void)hipGraphExecDestroy(ctx->sg_exec);
  if (ctx->sg_graph) (void)hipGraphDestroy(ctx->sg_graph);
  for (auto &r : ctx->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (void *p : ctx->allocs) (void)hipFree(p);
  if (ctx->helm_buf) (void)hipFree(ctx->helm_buf);
  if (ctx->mlp_buf) (void)hipFree(ctx->mlp_buf);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PHLX_OK;
}

extern "C" int phlx_nvars(const phlx_ctx *ctx) { return ctx ? ctx->g.nv : PHLX_E_ARG; }
extern "C" void *phlx_stream(phlx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ------------------------------------------------------------- host exchange

static int copy_in(phlx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return PHLX_OK;
}
static int copy_out(phlx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return PHLX_OK;
}

static int check_launch();
static int materialize(phlx_ctx *ctx);

// The energy predictor's per-zone records (GUESS_PRED) no longer describe
// the state (a new state, a diffusion step): NaN bytes make each zone start
// its next solve from the carried T.
static int forget_predictor(phlx_ctx *ctx) {
  if (!ctx->g.eprev) return PHLX_OK;
  HIPCHK(hipMemsetAsync(ctx->g.eprev, 0xFF, sizeof(double) * ctx->g.N, ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_upload_state(phlx_ctx *ctx, const double *prim, const double *eint,
                                 const double *temp, const double *b1, const double *b2,
                                 const double *b3) {
  if (!ctx) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  if (const int m = materialize(ctx)) return m;   // fields not uploaded stay current
  Grid &g = ctx->g;
  int rc = 0;
  if (prim) rc |= copy_in(ctx, g.P, prim, sizeof(double) * g.nv * g.N);
  if (eint) rc |= copy_in(ctx, g.eint, eint, sizeof(double) * g.N);
  // the ideal-gas EoS carries no temperature on the device; Helmholtz runs
  // take it as the warm start of the step-0 T(rho, p) solve
  if (temp && g.temp) rc |= copy_in(ctx, g.temp, temp, sizeof(double) * g.N);
  const double *bs[3] = {b1, b2, b3};
  for (int d = 0; d < g.sd; d++)
    if (bs[d]) {
      if (!g.mhd) return PHLX_E_ARG;
      rc |= copy_in(ctx, g.bf[d], bs[d], sizeof(double) * g.nf[d]);
    }
  if (rc) return PHLX_E_HIP;
  // mixed precision: the uploaded fields are the state now (U is not, and
  // nothing lags it); a state replaced between steps refreshes the sweeps'
  // FP32 inputs (before phlx_init, init does it)
  if (ctx->kt.prims) launch_flags(g.sc, 0, 0, ctx->stream);
  if ((rc = forget_predictor(ctx))) return rc;
  if (ctx->initialized && ctx->kt.pack) {
    ctx->kt.pack(ctx->g, ctx->stream);
    rc = check_launch();
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_download_state(phlx_ctx *ctx, double *prim, double *eint, double *temp,
                                   double *b1, double *b2, double *b3, double *bcc) {
  if (!ctx) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  if (const int m = materialize(ctx)) return m;
  Grid &g = ctx->g;
  int rc = 0;
  if (prim) rc |= copy_out(ctx, prim, g.P, sizeof(double) * g.nv * g.N);
  if (eint) rc |= copy_out(ctx, eint, g.eint, sizeof(double) * g.N);
  if (temp && g.temp) {
    rc |= copy_out(ctx, temp, g.temp, sizeof(double) * g.N);
  } else if (temp && g.tdiff && ctx->tdiff_valid) {
    // the ideal gas after a thermal-diffusion step: its eint_to_temp T
    rc |= copy_out(ctx, temp, g.tdiff, sizeof(double) * g.N);
  } else if (temp) {
    // T = p / (rho R / mu) for the ideal gas (source/source.F90:13670-13674)
    std::vector<double> p(2 * g.N);
    rc |= copy_out(ctx, p.data(), g.P, sizeof(double) * g.N);
    rc |= copy_out(ctx, p.data() + g.N, g.P + (size_t)g.ip * g.N, sizeof(double) * g.N);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (long long c = 0; c < g.N; c++) temp[c] = p[g.N + c] / (p[c] * kRgas * g.inv_mu);
  }
  double *bs[3] = {b1, b2, b3};
  for (int d = 0; d < g.sd; d++)
    if (bs[d] && g.mhd) rc |= copy_out(ctx, bs[d], g.bf[d], sizeof(double) * g.nf[d]);
  if (bcc) {
    if (g.mhd) {
      rc |= copy_out(ctx, bcc, g.bcc, sizeof(double) * g.sd * g.N);
      if (g.sd == 2) std::memset(bcc + 2 * g.N, 0, sizeof(double) * g.N);
    } else {
      std::memset(bcc, 0, sizeof(double) * 3 * g.N);
    }
  }
  if (rc) return PHLX_E_HIP;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

// Fortran AoS prim(1:nv, lx1-ngc:ux1+ngc, lx2-ngc:ux2+ngc, lx3-ngc:ux3+ngc)
// <-> interior SoA (source/source.F90:1219-1227).
static void aos_geom(const Grid &g, int ngc, long long &e1, long long &e2, long long &g3) {
  e1 = g.n1 + 2 * ngc;
  e2 = g.n2 + 2 * ngc;
  g3 = (g.sd == 3) ? ngc : 0;
}

extern "C" int phlx_upload_prim_aos(phlx_ctx *ctx, const double *prim, int32_t ngc) {
  if (!ctx || !prim || ngc < 0) return PHLX_E_ARG;
  const Grid &g = ctx->g;
  long long e1, e2, g3;
  aos_geom(g, ngc, e1, e2, g3);
  std::vector<double> soa((size_t)g.nv * g.N);
  for (int k = 0; k < g.n3; k++)
    for (int j = 0; j < g.n2; j++)
      for (int i = 0; i < g.n1; i++) {
        const long long a = (((k + g3) * e2 + (j + ngc)) * e1 + (i + ngc)) * g.nv;
        const long long c = ((long long)k * g.n2 + j) * g.n1 + i;
        for (int v = 0; v < g.nv; v++) soa[(size_t)v * g.N + c] = prim[a + v];
      }
  return phlx_upload_state(ctx, soa.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int phlx_download_prim_aos(phlx_ctx *ctx, double *prim, int32_t ngc) {
  if (!ctx || !prim || ngc < 0) return PHLX_E_ARG;
  const Grid &g = ctx->g;
  long long e1, e2, g3;
  aos_geom(g, ngc, e1, e2, g3);
  std::vector<double> soa((size_t)g.nv * g.N);
  int rc = phlx_download_state(ctx, soa.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr);
  if (rc) return rc;
  const long long e3 = g.n3 + 2 * g3;
  std::memset(prim, 0, sizeof(double) * g.nv * e1 * e2 * e3);
  for (int k = 0; k < g.n3; k++)
    for (int j = 0; j < g.n2; j++)
      for (int i = 0; i < g.n1; i++) {
        const long long a = (((k + g3) * e2 + (j + ngc)) * e1 + (i + ngc)) * g.nv;
        const long long c = ((long long)k * g.n2 + j) * g.n1 + i;
        for (int v = 0; v < g.nv; v++) prim[a + v] = soa[(size_t)v * g.N + c];
      }
  return PHLX_OK;
}

extern "C" int phlx_set_eos_profiles(phlx_ctx *ctx, const double *eq_gammae_cc,
                                     const double *eq_gammae_f) {
  if (!ctx) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  Grid &g = ctx->g;
  const size_t ny = g.n2;
  std::vector<double> h(2 * (ny + 1), 0.0);
  if (eq_gammae_cc) std::memcpy(h.data(), eq_gammae_cc, ny * sizeof(double));
  if (eq_gammae_f) std::memcpy(h.data() + (ny + 1), eq_gammae_f, (ny + 1) * sizeof(double));
  HIPCHK(hipMemcpyAsync((void *)g.eq_ge_cc, h.data(), h.size() * sizeof(double),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_download_gammaf(phlx_ctx *ctx, double *gammaf) {
  if (!ctx || !gammaf) return PHLX_E_ARG;
  if (!ctx->g.gam) return PHLX_E_STATE;
  HIPCHK(hipSetDevice(ctx->device));
  if (const int m = materialize(ctx)) return m;
  HIPCHK(hipMemcpyAsync(gammaf, ctx->g.gam, sizeof(double) * 2 * ctx->g.N,
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_set_profiles(phlx_ctx *ctx, const double *eq_rho_cc, const double *eq_p_cc,
                                 const double *eq_rho_f, const double *eq_p_f,
                                 const double *grav_cc, const double *phi_cc,
                                 const double *phi_f, const double *edot_cc) {
  if (!ctx) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  Grid &g = ctx->g;
  const size_t ny = g.n2;
  std::vector<double> h(8 * (ny + 1), 0.0);
  const double *src[8] = {eq_rho_cc, eq_p_cc, eq_rho_f, eq_p_f, grav_cc, phi_cc, phi_f, edot_cc};
  const size_t len[8] = {ny, ny, ny + 1, ny + 1, ny, ny, ny + 1, ny};
  for (int a = 0; a < 8; a++)
    if (src[a]) std::memcpy(h.data() + a * (ny + 1), src[a], len[a] * sizeof(double));
  HIPCHK(hipMemcpyAsync((void *)g.eq_rho_cc, h.data(), h.size() * sizeof(double),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

// ------------------------------------------------------------- profiling

static int kbegin(phlx_ctx *ctx, int kid, ProfRec &r) {
  if (!ctx->prof) return 0;
  r.kid = kid;
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
  (void)hipEventRecord(r.a, ctx->stream);
  return 1;
}
static void kend(phlx_ctx *ctx, int on, ProfRec &r) {
  if (on != 1) return;
  (void)hipEventRecord(r.b, ctx->stream);
  ctx->recs.push_back(r);
}

static int harvest(phlx_ctx *ctx) {
  if (ctx->recs.empty()) return PHLX_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto &r : ctx->recs) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    ctx->prof_ms[r.kid] += ms;
    ctx->prof_n[r.kid] += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  ctx->recs.clear();
  return PHLX_OK;
}

extern "C" int phlx_profile_enable(phlx_ctx *ctx, int32_t on) {
  if (!ctx) return PHLX_E_ARG;
  int rc = harvest(ctx);
  if (rc) return rc;
  ctx->prof = on != 0;
  for (int i = 0; i < KID_COUNT; i++) {
    ctx->prof_ms[i] = 0.0;
    ctx->prof_n[i] = 0;
  }
  return PHLX_OK;
}

extern "C" int phlx_profile_read(phlx_ctx *ctx, int32_t kernel, double *total_ms,
                                 int64_t *launches) {
  if (!ctx || kernel < 0 || kernel >= KID_COUNT) return PHLX_E_ARG;
  int rc = harvest(ctx);
  if (rc) return rc;
  if (total_ms) *total_ms = ctx->prof_ms[kernel];
  if (launches) *launches = ctx->prof_n[kernel];
  return PHLX_OK;
}

// ------------------------------------------------------------- stepping

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::fprintf(stderr, "phlx: launch failed: %s\n", hipGetErrorString(e));
    return PHLX_E_HIP;
  }
  return PHLX_OK;
}

// The surrogate (its wave-wide inference), the bracketed start and the
// iteration counters run in the guess variant of the update; the carried-T
// and predictor starts in the plain one.
static bool needs_guess_variant(int mode, bool count) {
  return count || mode == GUESS_MLP || mode == GUESS_BRACKET;
}

extern "C" int phlx_stage_begin(phlx_ctx *ctx, int32_t irk) {
  // state0 / b0 capture is fused into the first stage's update and face-B
  // kernels; the stage starts with the shock-flattening flags of the
  // current state (USE_SHOCK_FLATTENING), else launches nothing.
  if (!ctx) return PHLX_E_ARG;
  if (irk < 0 || irk >= ctx->stages) return PHLX_E_ARG;
  if (!ctx->initialized) return PHLX_E_STATE;
  if (ctx->g.sflat) {
    ProfRec r;
    int on = kbegin(ctx, KID_MISC, r);
    launch_shock_flags(ctx->g, ctx->stream);
    kend(ctx, on, r);
    return check_launch();
  }
  return PHLX_OK;
}

extern "C" int phlx_sweep(phlx_ctx *ctx, int32_t dir) {
  if (!ctx || dir < 0 || dir >= ctx->g.sd) return PHLX_E_ARG;
  ProfRec r;
  int on = kbegin(ctx, KID_FLUX1 + dir, r);
  ctx->kt.flux(ctx->g, dir, ctx->stream);
  kend(ctx, on, r);
  return check_launch();
}

extern "C" int phlx_ct(phlx_ctx *ctx, int32_t irk, double a1, double a2, double a3) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->g.mhd) return PHLX_OK;
  // RK2: stage 1 writes the new field into the b0f buffers (bf keeps the
  // step-start field: no copy of it) and the pointer sets swap, so every
  // later launch reads the stage-1 field through g.bf; stage 2 writes its
  // field over the step-start one and the sets swap back.  RK3 keeps the
  // step-start copy.
  int mode = irk == 0 ? FACEB_FIRST : FACEB_LATER;
  if (ctx->stages == 2) {
    if ((irk == 0) == ctx->b_swapped) return PHLX_E_STATE;  // stage order
    mode = irk == 0 ? FACEB_RK2_1 : FACEB_RK2_2;
  }
  ProfRec r;
  int on = kbegin(ctx, KID_CORNER, r);
  ctx->kt.corner(ctx->g, ctx->stream);
  kend(ctx, on, r);
  on = kbegin(ctx, KID_FACEB, r);
  ctx->kt.faceb(ctx->g, mode, a1, a2, a3, ctx->stream);
  kend(ctx, on, r);
  if (ctx->stages == 2) {
    for (int d = 0; d < 3; d++) std::swap(ctx->g.bf[d], ctx->g.b0f[d]);
    ctx->b_swapped = !ctx->b_swapped;
  }
  return check_launch();
}

// swept: the fused sweep formed the stage's U; finish it (cons -> prim, EoS,
// CFL) instead of the flux-divergence update.
static int rk_update(phlx_ctx *ctx, int32_t irk, double a1, double a2, double a3,
                     int32_t last_stage, bool swept) {
  ctx->tdiff_valid = false;
  ProfRec r;
  int on = kbegin(ctx, KID_UPDATE, r);
  const bool guess = ctx->kt.update_guess && needs_guess_variant(ctx->g.eos_guess, ctx->eos_stats);
  // stage 1 of the mixed path starts from U when the device flags say it is
  // current (the FP64 conserved state of the last step that ran); a last
  // stage that runs sets them, so a step past tmax leaves them unchanged
  const int first = irk == 0 ? 1 : 0;
  const int fin = irk == ctx->stages - 1;
  const int last = last_stage != 0;
  if (swept && guess)
    ctx->kt.update_swept_guess(ctx->g, first, last, a1, a2, a3, fin, ctx->stream);
  else if (swept)
    ctx->kt.update_swept(ctx->g, first, last, a1, a2, a3, fin, ctx->stream);
  else if (guess)
    ctx->kt.update_guess(ctx->g, first, last, a1, a2, a3, fin, ctx->stream);
  else
    ctx->kt.update(ctx->g, first, last, a1, a2, a3, fin, ctx->stream);
  if (fin && ctx->kt.prims) ctx->prim_stale = true;
  kend(ctx, on, r);
  return check_launch();
}

extern "C" int phlx_rk_update(phlx_ctx *ctx, int32_t irk, double a1, double a2, double a3,
                              int32_t last_stage) {
  if (!ctx) return PHLX_E_ARG;
  return rk_update(ctx, irk, a1, a2, a3, last_stage, false);
}

// Whether hydro_step takes the fused sweep: a family that has one, no shock
// flattening (its flagged cells switch scheme per face), and x1 tiles that
// divide the grid.
static bool use_fused_sweep(const phlx_ctx *ctx) {
  const Grid &g = ctx->g;
  // measured on MI355X: without MHD the fused stage is 5-19 % faster (FP64
  // and mixed); with MHD the split sweeps are 1-5 % faster (FP32: the
  // column segments prefetch a cell ahead at 3-4 waves/SIMD; FP64: the
  // fused MHD sweep spills)
  const bool want = ctx->fused_sweep < 0 ? !ctx->g.mhd : ctx->fused_sweep != 0;
  return want && ctx->kt.sweep && !g.sflat && g.sd == 3 && g.n1 % kSweepT == 0 &&
         (!ctx->kt.update_guess || ctx->kt.update_swept_guess);
}

extern "C" int phlx_set_fused_sweep(phlx_ctx *ctx, int32_t mode) {
  if (!ctx || mode < -1 || mode > 1) return PHLX_E_ARG;
  ctx->fused_sweep = mode;
  return PHLX_OK;
}

static int diffusion_step(phlx_ctx *ctx, int strang);
static int parabolic_dt(phlx_ctx *ctx);

// Brings the FP64 prim / eint / b_cc of the mixed path up to date with U
// (a no-op in FP64 runs and when no step was launched since the last call;
// on the device the rebuild runs only if a step actually ran, SC_P_STALE).
static int materialize(phlx_ctx *ctx) {
  if (!ctx->prim_stale) return PHLX_OK;
  ctx->kt.prims(ctx->g, ctx->stream);
  launch_flags(ctx->g.sc, ~SC_P_STALE, 0, ctx->stream);
  ctx->prim_stale = false;
  return check_launch();
}

static int do_hydro_step(phlx_ctx *ctx, bool with_cfl) {
  const bool fused = use_fused_sweep(ctx);
  for (int irk = 0; irk < ctx->stages; irk++) {
    int rc = phlx_stage_begin(ctx, irk);
    if (rc) return rc;
    const double a1 = ctx->rk[irk][0], a2 = ctx->rk[irk][1], a3 = ctx->rk[irk][2];
    if (fused) {
      ProfRec r;
      int on = kbegin(ctx, KID_SWEEP, r);
      ctx->kt.sweep(ctx->g, irk == 0 ? 1 : 0, irk == ctx->stages - 1 ? 1 : 0, a1, a2, a3,
                    ctx->stream);
      kend(ctx, on, r);
      rc = check_launch();
      if (rc) return rc;
    } else {
      for (int d = 0; d < ctx->g.sd; d++) {
        rc = phlx_sweep(ctx, d);
        if (rc) return rc;
      }
    }
    rc = phlx_ct(ctx, irk, a1, a2, a3);
    if (rc) return rc;
    rc = rk_update(ctx, irk, a1, a2, a3, with_cfl && irk == ctx->stages - 1, fused);
    if (rc) return rc;
  }
  return PHLX_OK;
}

extern "C" int phlx_init(phlx_ctx *ctx) {
  if (!ctx) return PHLX_E_ARG;
  if (ctx->cfg.eos != PHLX_EOS_IDEAL && !ctx->helm_loaded) return PHLX_E_TABLE;
  HIPCHK(hipSetDevice(ctx->device));
  Scalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.tmax = ctx->cfg.tmax;
  s0.cfl = ctx->cfg.cfl;
  s0.active = 1;
  HIPCHK(hipMemcpyAsync(ctx->g.sc, &s0, sizeof(s0), hipMemcpyHostToDevice, ctx->stream));
  if (const int m = materialize(ctx)) return m;
  ctx->b_swapped = false;  // g.bf is the current face field whichever buffer it is
  ctx->kt.init(ctx->g, 0, ctx->stream);
  if (ctx->kt.pack) ctx->kt.pack(ctx->g, ctx->stream);
  int rc = check_launch();
  if (rc) return rc;
  ctx->initialized = true;
  ctx->tdiff_valid = false;
  // without EVALUATE_PARABOLIC_TIMESTEP dt_par is computed once, at step 0
  // (source/source.F90:6504-6512)
  if (ctx->diff_on && !ctx->diff_eval && (rc = parabolic_dt(ctx))) return rc;
  return phlx_compute_dt(ctx, nullptr, nullptr);
}

extern "C" int phlx_resume(phlx_ctx *ctx, double time, double dt, double smax, int64_t step) {
  if (!ctx || !(dt > 0.0) || !(smax > 0.0) || step < 0) return PHLX_E_ARG;
  if (ctx->cfg.eos != PHLX_EOS_IDEAL && !ctx->helm_loaded) return PHLX_E_TABLE;
  HIPCHK(hipSetDevice(ctx->device));
  Scalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.time = time;
  s0.dt = dt;
  s0.smax = smax;
  s0.step = step;
  s0.tmax = ctx->cfg.tmax;
  s0.cfl = ctx->cfg.cfl;
  s0.active = time < ctx->cfg.tmax ? 1 : 0;
  HIPCHK(hipMemcpyAsync(ctx->g.sc, &s0, sizeof(s0), hipMemcpyHostToDevice, ctx->stream));
  if (const int m = materialize(ctx)) return m;
  ctx->b_swapped = false;
  ctx->kt.init(ctx->g, 1, ctx->stream);
  if (ctx->kt.pack) ctx->kt.pack(ctx->g, ctx->stream);
  int rc = check_launch();
  if (rc) return rc;
  ctx->initialized = true;
  return PHLX_OK;
}

extern "C" int phlx_compute_dt(phlx_ctx *ctx, double *dt_out, double *smax_out) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->initialized) return PHLX_E_STATE;
  if (const int m = materialize(ctx)) return m;
  ctx->kt.cfl(ctx->g, ctx->stream);
  launch_step_end(ctx->g, 0, ctx->stream);
  int rc = check_launch();
  if (rc) return rc;
  if (dt_out || smax_out) {
    double t, dt, smax;
    int64_t st;
    rc = phlx_get_scalars(ctx, &t, &dt, &smax, &st);
    if (rc) return rc;
    if (dt_out) *dt_out = dt;
    if (smax_out) *smax_out = smax;
  }
  return PHLX_OK;
}

extern "C" int phlx_hydro_step(phlx_ctx *ctx) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->initialized) return PHLX_E_STATE;
  HIPCHK(hipSetDevice(ctx->device));
  return do_hydro_step(ctx, false);
}

extern "C" int phlx_set_step_graph(phlx_ctx *ctx, int32_t on) {
  if (!ctx || on < 0 || on > 1) return PHLX_E_ARG;
  ctx->step_graph = on != 0;
  return PHLX_OK;
}

static void drop_step_graph(phlx_ctx *ctx) {
  if (ctx->sg_exec) (void)hipGraphExecDestroy(ctx->sg_exec);
  if (ctx->sg_graph) (void)hipGraphDestroy(ctx->sg_graph);
  ctx->sg_exec = nullptr;
  ctx->sg_graph = nullptr;
}

// The executable graph of one step (step begin, hydro_step with the CFL,
// step end), captured from the same launch sequence as the eager loop.  It
// holds only kernel nodes.  Captured again when the grid descriptor, the
// sweep form or the Newton counting differ from the capture's.
static int step_graph_ready(phlx_ctx *ctx) {
  const bool fused = use_fused_sweep(ctx);
  if (ctx->sg_exec && ctx->sg_fused == fused && ctx->sg_stats == ctx->eos_stats &&
      std::memcmp(&ctx->sg_g, &ctx->g, sizeof(Grid)) == 0)
    return PHLX_OK;
  drop_step_graph(ctx);
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  launch_step_begin(ctx->g.sc, ctx->stream);
  const int rc = do_hydro_step(ctx, true);
  launch_step_end(ctx->g, 1, ctx->stream);
  hipGraph_t gr = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
  if (rc || e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
    return rc ? rc : PHLX_E_HIP;
  }
  ctx->sg_graph = gr;
  if (hipGraphInstantiate(&ctx->sg_exec, gr, nullptr, nullptr, 0) != hipSuccess) {
    ctx->sg_exec = nullptr;
    drop_step_graph(ctx);
    return PHLX_E_HIP;
  }
  ctx->sg_g = ctx->g;
  ctx->sg_fused = fused;
  ctx->sg_stats = ctx->eos_stats;
  return PHLX_OK;
}

extern "C" int phlx_run(phlx_ctx *ctx, int32_t nsteps) {
  if (!ctx || nsteps < 0) return PHLX_E_ARG;
  if (!ctx->initialized) return PHLX_E_STATE;
  HIPCHK(hipSetDevice(ctx->device));
  const bool diff = ctx->diff_on;
  if (ctx->step_graph && nsteps > 0 && !diff && !ctx->prof && !ctx->b_swapped) {
    const int rc = step_graph_ready(ctx);
    if (rc) return rc;
    for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(ctx->sg_exec, ctx->stream));
    // the host-side records an eager step leaves
    ctx->tdiff_valid = false;
    if (ctx->kt.prims) ctx->prim_stale = true;
    return PHLX_OK;
  }
  for (int s = 0; s < nsteps; s++) {
    launch_step_begin(ctx->g.sc, ctx->stream);
    int rc;
    if (diff) {
      // THERMAL_DIFFUSION_STS around hydro_step (source/source.F90:6672-6688);
      // the hyperbolic dt follows the second half (:6716)
      if (ctx->diff_eval && (rc = parabolic_dt(ctx))) return rc;
      if ((rc = diffusion_step(ctx, 1))) return rc;
    }
    rc = do_hydro_step(ctx, !diff);
    if (rc) return rc;
    if (diff) {
      if ((rc = diffusion_step(ctx, 2))) return rc;
      ctx->kt.cfl(ctx->g, ctx->stream);
    }
    launch_step_end(ctx->g, 1, ctx->stream);
    rc = check_launch();
    if (rc) return rc;
  }
  return PHLX_OK;
}

extern "C" int phlx_synchronize(phlx_ctx *ctx) {
  if (!ctx) return PHLX_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_get_scalars(phlx_ctx *ctx, double *time, double *dt, double *smax,
                                int64_t *step) {
  if (!ctx) return PHLX_E_ARG;
  Scalars s;
  HIPCHK(hipMemcpyAsync(&s, ctx->g.sc, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (time) *time = s.time;
  if (dt) *dt = s.dt;
  if (smax) *smax = s.smax;
  if (step) *step = s.step;
  return PHLX_OK;
}

extern "C" int phlx_set_scalars(phlx_ctx *ctx, double time, double dt, double smax,
                                int64_t step) {
  if (!ctx) return PHLX_E_ARG;
  Scalars s;
  HIPCHK(hipMemcpyAsync(&s, ctx->g.sc, sizeof(s), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  s.time = time;
  s.dt = dt;
  s.smax = smax;
  s.step = step;
  HIPCHK(hipMemcpyAsync(ctx->g.sc, &s, sizeof(s), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return PHLX_OK;
}

extern "C" int phlx_divb_max(phlx_ctx *ctx, double *out) {
  if (!ctx || !out) return PHLX_E_ARG;
  if (!ctx->g.mhd) {
    *out = 0.0;
    return PHLX_OK;
  }
  ctx->kt.divb(ctx->g, ctx->stream);
  launch_reduce_to(ctx->g, ctx->scratch_val, ctx->stream);
  int rc = check_launch();
  if (rc) return rc;
  double v = 0.0;
  HIPCHK(hipMemcpyAsync(&v, ctx->scratch_val, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = v;
  return PHLX_OK;
}

extern "C" int phlx_stream_copy(phlx_ctx *ctx, int64_t nbytes, int32_t reps, double *gbps) {
  if (!ctx || nbytes < 16 || reps < 1 || !gbps) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const long long n16 = nbytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return PHLX_E_NOMEM;
  if (hipMalloc(&b, n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    return PHLX_E_NOMEM;
  }
  int rc = PHLX_OK;
  hipEvent_t e0, e1;
  if (hipMemsetAsync(a, 0, n16 * 16, ctx->stream) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = PHLX_E_HIP;
  } else {
    launch_stream_copy(a, b, n16, ctx->stream);  // warm
    (void)hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < reps; r++) launch_stream_copy(a, b, n16, ctx->stream);
    (void)hipEventRecord(e1, ctx->stream);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = PHLX_E_HIP;
    else
      *gbps = 2.0 * (double)n16 * 16.0 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// ------------------------------------------------------------- Helmholtz EoS

static int load_table(phlx_ctx *ctx, const double *f9, const double *c4, int32_t nrho,
                      int32_t nT, double ldlo, double ldhi, double ltlo, double lthi, int pig);

extern "C" int phlx_load_helm_table(phlx_ctx *ctx, const double *f9, const double *c4,
                                    int32_t nrho, int32_t nT, double ldlo, double ldhi,
                                    double ltlo, double lthi) {
  if (ctx && ctx->cfg.eos == PHLX_EOS_PIG) return PHLX_E_CONFIG;
  return load_table(ctx, f9, c4, nrho, nT, ldlo, ldhi, ltlo, lthi, 0);
}

// load_pig_table (source/source.F90:25407-25455): the same file columns.
extern "C" int phlx_load_pig_table(phlx_ctx *ctx, const double *f9, const double *c4,
                                   int32_t nrho, int32_t nT, double ldlo, double ldhi,
                                   double ltlo, double lthi) {
  if (ctx && ctx->cfg.eos == PHLX_EOS_HELMHOLTZ) return PHLX_E_CONFIG;
  return load_table(ctx, f9, c4, nrho, nT, ldlo, ldhi, ltlo, lthi, 1);
}

static int load_table(phlx_ctx *ctx, const double *f9, const double *c4, int32_t nrho,
                      int32_t nT, double ldlo, double ldhi, double ltlo, double lthi, int pig) {
  if (!ctx || !f9 || !c4) return PHLX_E_ARG;
  if (nrho < 2 || nT < 2 || !(ldhi > ldlo) || !(lthi > ltlo)) return PHLX_E_TABLE;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const size_t np = (size_t)nrho * nT;
  for (size_t p = 0; p < 9 * np; p++)
    if (!std::isfinite(f9[p])) return PHLX_E_TABLE;
  for (size_t p = 0; p < 4 * np; p++)
    if (!std::isfinite(c4[p])) return PHLX_E_TABLE;
  // node values computed on the host exactly as the reference builds its
  // axes (source/source.F90:23094-23104)
  HelmDev h{};
  h.nr = nrho;
  h.nt = nT;
  h.dt = (lthi - ltlo) / (double)(nT - 1);
  h.dr = (ldhi - ldlo) / (double)(nrho - 1);
  std::vector<double> nodes((size_t)nrho + nT);
  for (int i = 0; i < nrho; i++) nodes[i] = std::pow(10.0, ldlo + (double)i * h.dr);
  for (int j = 0; j < nT; j++) nodes[nrho + j] = std::pow(10.0, ltlo + (double)j * h.dt);
  h.tl = std::pow(10.0, ltlo);
  h.th = std::pow(10.0, lthi);
  h.lrho1 = std::log10(nodes[0]);
  h.lT1 = std::log10(nodes[nrho]);
  // one allocation: records | nodes | staging of the file-order blocks
  const size_t rec_b = sizeof(double) * kHelmRec * np;
  const size_t node_b = sizeof(double) * nodes.size();
  const size_t stage_b = sizeof(double) * 13 * np;
  void *buf = nullptr;
  if (hipMalloc(&buf, rec_b + node_b + stage_b + 512) != hipSuccess) return PHLX_E_NOMEM;
  char *base = (char *)buf;
  double *tab = (double *)base;
  double *nd = (double *)(base + rec_b);
  double *st9 = (double *)(base + ((rec_b + node_b + 255) / 256) * 256);
  double *st4 = st9 + 9 * np;
  int rc = PHLX_OK;
  if (hipMemcpyAsync(nd, nodes.data(), node_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(st9, f9, sizeof(double) * 9 * np, hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess ||
      hipMemcpyAsync(st4, c4, sizeof(double) * 4 * np, hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess) {
    rc = PHLX_E_HIP;
  } else {
    launch_helm_pack((long long)np, st9, st4, tab, ctx->stream);
    rc = check_launch();
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PHLX_E_HIP;
  }
  if (rc) {
    (void)hipFree(buf);
    return rc;
  }
  if (ctx->helm_buf) (void)hipFree(ctx->helm_buf);
  ctx->helm_buf = buf;
  h.tab = tab;
  h.rho = nd;
  h.T = nd + nrho;
  h.pig = pig;
  ctx->helm = h;
  ctx->g.helm = h;
  ctx->helm_loaded = true;
  return PHLX_OK;
}

extern "C" int phlx_helm_rhoT_full_batch(phlx_ctx *ctx, int64_t n, const double *rho,
                                         const double *T, const double *abar,
                                         const double *zbar, double *P, double *E,
                                         double *sound, double *cv) {
  if (!ctx || n < 0) return PHLX_E_ARG;
  if (!ctx->helm_loaded) return PHLX_E_TABLE;
  if (n == 0) return PHLX_OK;
  if (!rho || !T || !abar || !zbar) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  launch_helm_full(ctx->helm, n, rho, T, abar, zbar, P, E, sound, cv, ctx->stream);
  return check_launch();
}

extern "C" int phlx_helm_rhoe_given_batch(phlx_ctx *ctx, int64_t n, const double *rho,
                                          const double *e, const double *abar,
                                          const double *zbar, double *T, double *P,
                                          double *sound, int32_t return_sound, int32_t *iters) {
  if (!ctx || n < 0) return PHLX_E_ARG;
  if (!ctx->helm_loaded) return PHLX_E_TABLE;
  if (n == 0) return PHLX_OK;
  if (!rho || !e || !abar || !zbar || !T) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  launch_helm_rhoe(ctx->helm, n, rho, e, abar, zbar, T, P, sound, return_sound, iters,
                   ctx->stream);
  return check_launch();
}

extern "C" int phlx_helm_rhoP_given_batch(phlx_ctx *ctx, int64_t n, const double *rho,
                                          const double *p, const double *abar,
                                          const double *zbar, double *T, double *E,
                                          double *sound, int32_t return_sound, int32_t *iters) {
  if (!ctx || n < 0) return PHLX_E_ARG;
  if (!ctx->helm_loaded) return PHLX_E_TABLE;
  if (n == 0) return PHLX_OK;
  if (!rho || !p || !abar || !zbar || !T) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  launch_helm_rhoP(ctx->helm, n, rho, p, abar, zbar, T, E, sound, return_sound, iters,
                   ctx->stream);
  return check_launch();
}

// ------------------------------------------------------------- T(rho, e) surrogate

constexpr int kMlpMaxGradBlocks = 1024;  // gradient-check batches up to 131072 samples

static int mlp_alloc(phlx_ctx *ctx) {
  if (ctx->mlp_buf) return PHLX_OK;
  const size_t P = kMlpNParam;
  const size_t floats =
      3 * P + 256 + 3 * 64 + (size_t)kMlpMaxGradBlocks * P + kMlpMaxGradBlocks + P;
  const size_t bytes =
      floats * sizeof(float) + (kMlpW2aShorts + kMlpW1aShorts) * sizeof(unsigned short) + 4096;
  void *buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) return PHLX_E_NOMEM;
  HIPCHK(hipMemsetAsync(buf, 0, bytes, ctx->stream));
  float *f = (float *)buf;
  ctx->mlp_theta = f; f += P;
  ctx->mlp_m = f; f += P;
  ctx->mlp_v = f; f += P;
  ctx->mlp_w1q = f; f += 256;
  ctx->mlp_w3q = f; f += 3 * 64;  // w3q, then the inference's scaled b1, b2
  ctx->mlp_gpart = f; f += (size_t)kMlpMaxGradBlocks * P;
  ctx->mlp_lpart = f; f += kMlpMaxGradBlocks;
  ctx->mlp_gsum = f; f += P;
  uintptr_t a = ((uintptr_t)f + 255) & ~(uintptr_t)255;
  ctx->mlp_wa = (unsigned short *)a;
  ctx->mlp_buf = buf;
  MlpDev &m = ctx->g.mlp;
  m.theta = ctx->mlp_theta;
  m.w1q = ctx->mlp_w1q;
  m.w3q = ctx->mlp_w3q;
  m.bs = ctx->mlp_w3q + kMlpH;
  m.w2a = (const mlp_bf8 *)ctx->mlp_wa;
  m.w1a = (const mlp_s4 *)(ctx->mlp_wa + kMlpW2aShorts);
  return PHLX_OK;
}

static void mlp_set_norm(MlpDev &m, const float *norm) {
  for (int c = 0; c < 4; c++) {
    m.in_off[c] = norm[c];
    m.in_scale[c] = norm[4 + c];
  }
  m.out_off = norm[8];
  m.out_scale = norm[9];
}

extern "C" int phlx_mlp_set_params(phlx_ctx *ctx, const float *theta, const float *norm) {
  if (!ctx || !theta || !norm) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = mlp_alloc(ctx);
  if (rc) return rc;
  for (int i = 0; i < kMlpNParam; i++)
    if (!std::isfinite(theta[i])) return PHLX_E_ARG;
  if (!(norm[9] > 0.0f)) return PHLX_E_ARG;
  HIPCHK(hipMemcpyAsync(ctx->mlp_theta, theta, sizeof(float) * kMlpNParam, hipMemcpyHostToDevice,
                        ctx->stream));
  launch_mlp_quantize(ctx->mlp_theta, ctx->mlp_w1q, ctx->mlp_w3q, ctx->mlp_wa, ctx->stream);
  rc = check_launch();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  mlp_set_norm(ctx->g.mlp, norm);
  ctx->mlp_ready = true;
  return PHLX_OK;
}

extern "C" int phlx_mlp_get_params(phlx_ctx *ctx, float *theta, float *norm) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->mlp_ready) return PHLX_E_STATE;
  if (theta) {
    HIPCHK(hipMemcpyAsync(theta, ctx->mlp_theta, sizeof(float) * kMlpNParam,
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (norm) {
    const MlpDev &m = ctx->g.mlp;
    for (int c = 0; c < 4; c++) {
      norm[c] = m.in_off[c];
      norm[4 + c] = m.in_scale[c];
    }
    norm[8] = m.out_off;
    norm[9] = m.out_scale;
  }
  return PHLX_OK;
}

extern "C" int phlx_mlp_forward(phlx_ctx *ctx, int64_t n, const float *raw, float *y,
                                int32_t *inbox) {
  if (!ctx || n < 0) return PHLX_E_ARG;
  if (!ctx->mlp_ready) return PHLX_E_STATE;
  if (n == 0) return PHLX_OK;
  if (!raw || !y) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  launch_mlp_forward(ctx->g.mlp, n, raw, y, inbox, ctx->stream);
  return check_launch();
}

extern "C" int phlx_mlp_grad(phlx_ctx *ctx, int64_t n, const float *xn, const float *tn,
                             const float *wt, float *grad, double *loss) {
  if (!ctx || n <= 0 || !xn || !tn || !grad) return PHLX_E_ARG;
  if (!ctx->mlp_ready) return PHLX_E_STATE;
  if (n > (int64_t)kMlpMaxGradBlocks * 128) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const int nblk = launch_mlp_grad(ctx->mlp_theta, ctx->mlp_w1q, ctx->mlp_w3q, xn, tn, wt, 0, n,
                                   ctx->mlp_gpart, ctx->mlp_lpart, ctx->stream);
  launch_mlp_adam(ctx->mlp_theta, ctx->mlp_m, ctx->mlp_v, ctx->mlp_gpart, nblk, 1, 0.0f,
                  ctx->mlp_w1q, ctx->mlp_w3q, ctx->mlp_wa, grad, ctx->stream);
  int rc = check_launch();
  if (rc) return rc;
  if (loss) {
    std::vector<float> lp(nblk);
    HIPCHK(hipMemcpyAsync(lp.data(), ctx->mlp_lpart, sizeof(float) * nblk, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = s / (double)n;
  }
  return PHLX_OK;
}

extern "C" int phlx_mlp_set_train_box(phlx_ctx *ctx, const double *box) {
  if (!ctx) return PHLX_E_ARG;
  if (!box) {
    ctx->mlp_has_box = false;
    return PHLX_OK;
  }
  for (int c = 0; c < 8; c += 2)
    if (!(box[c + 1] > box[c]) || !std::isfinite(box[c]) || !std::isfinite(box[c + 1]))
      return PHLX_E_ARG;
  if (!(box[4] > 0.0) || !(box[6] > 0.0) || box[7] > 1.0) return PHLX_E_ARG;
  for (int c = 0; c < 8; c++) ctx->mlp_box[c] = box[c];
  ctx->mlp_has_box = true;
  return PHLX_OK;
}

extern "C" int phlx_mlp_train(phlx_ctx *ctx, int64_t nsamples, int32_t epochs, int32_t batch,
                              double lr, uint64_t seed, double *train_loss, double *val_rms_logT,
                              double *val_rms_loge) {
  if (!ctx || nsamples < 4096 || epochs < 1 || batch < 128 || batch % 128 ||
      batch > kMlpMaxGradBlocks * 128 || !(lr > 0.0))
    return PHLX_E_ARG;
  if (!ctx->helm_loaded) return PHLX_E_TABLE;
  if (ctx->helm.pig) return PHLX_E_CONFIG;  // the surrogate's inputs are Helmholtz (abar, zbar)
  HIPCHK(hipSetDevice(ctx->device));
  int rc = mlp_alloc(ctx);
  if (rc) return rc;
  const HelmDev &h = ctx->helm;
  const long long n = nsamples, nval = n / 16, ntr = n - nval;
  if (ntr < batch) return PHLX_E_ARG;
  // sample box: the table's rho*Ye and T ranges (0.05 dex inside the edges),
  // abar in [4, 16], Ye in [0.45, 0.5]; a box set by phlx_mlp_set_train_box
  // replaces it, clipped to the same table range
  const double ldlo = h.lrho1 + 0.05, ldhi = h.lrho1 + h.dr * (h.nr - 1) - 0.05;
  const double ltlo = h.lT1 + 0.05, lthi = h.lT1 + h.dt * (h.nt - 1) - 0.05;
  double box[8] = {ldlo, ldhi, ltlo, lthi, 4.0, 16.0, 0.45, 0.5};
  if (ctx->mlp_has_box) {
    for (int c = 0; c < 8; c++) box[c] = ctx->mlp_box[c];
    box[0] = std::max(box[0], ldlo);
    box[1] = std::min(box[1], ldhi);
    box[2] = std::max(box[2], ltlo);
    box[3] = std::min(box[3], lthi);
    if (!(box[1] > box[0]) || !(box[3] > box[2])) return PHLX_E_ARG;
  }
  float *data = nullptr;
  if (hipMalloc(&data, sizeof(float) * 11 * (size_t)n) != hipSuccess) return PHLX_E_NOMEM;
  float *raw = data, *label = data + 4 * n, *xn = data + 5 * n, *tn = data + 9 * n,
        *wt = data + 10 * n;
  launch_mlp_data(h, n, seed, box, raw, label, wt, ctx->stream);
  rc = check_launch();
  std::vector<float> hraw;
  if (!rc) {
    hraw.resize(4 * (size_t)n);
    if (hipMemcpyAsync(hraw.data(), raw, sizeof(float) * 4 * n, hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = PHLX_E_HIP;
  }
  if (rc) {
    (void)hipFree(data);
    return rc;
  }
  float norm[10];
  for (int c = 0; c < 4; c++) {
    float lo = hraw[c], hi = hraw[c];
    for (long long i = 1; i < n; i++) {
      lo = std::min(lo, hraw[4 * i + c]);
      hi = std::max(hi, hraw[4 * i + c]);
    }
    norm[c] = 0.5f * (lo + hi);
    norm[4 + c] = hi > lo ? 2.0f / (hi - lo) : 1.0f;
  }
  norm[8] = (float)(0.5 * (box[2] + box[3]));
  norm[9] = (float)(0.5 * (box[3] - box[2]));
  mlp_set_norm(ctx->g.mlp, norm);
  launch_mlp_normalize(ctx->g.mlp, n, raw, label, xn, tn, ctx->stream);
  // Xavier-uniform initial weights (host RNG), zero biases and moments
  std::vector<float> th(kMlpNParam, 0.0f);
  unsigned long long st = seed ^ 0x5DEECE66Dull;
  auto next = [&st]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = std::sqrt(6.0 / (4 + 64)), a2 = std::sqrt(6.0 / (64 + 64)),
               a3 = std::sqrt(6.0 / (64 + 1));
  for (int i = 0; i < 256; i++) th[kMlpOffW1 + i] = (float)((2.0 * next() - 1.0) * a1);
  for (int i = 0; i < 4096; i++) th[kMlpOffW2 + i] = (float)((2.0 * next() - 1.0) * a2);
  for (int i = 0; i < 64; i++) th[kMlpOffW3 + i] = (float)((2.0 * next() - 1.0) * a3);
  if (hipMemcpyAsync(ctx->mlp_theta, th.data(), sizeof(float) * kMlpNParam, hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->mlp_m, 0, sizeof(float) * 2 * kMlpNParam, ctx->stream) != hipSuccess) {
    (void)hipFree(data);
    return PHLX_E_HIP;
  }
  launch_mlp_quantize(ctx->mlp_theta, ctx->mlp_w1q, ctx->mlp_w3q, ctx->mlp_wa, ctx->stream);
  const long long spe = ntr / batch;  // steps per epoch
  const long long steps = spe * epochs;
  int nblk = 1;
  for (long long t = 1; t <= steps; t++) {
    const long long n0 = ((t - 1) % spe) * batch;
    const double cosf_ = 0.5 * (1.0 + std::cos(3.141592653589793 * (double)(t - 1) / (double)steps));
    const float lr_t = (float)(lr * (0.05 + 0.95 * cosf_));
    nblk = launch_mlp_grad(ctx->mlp_theta, ctx->mlp_w1q, ctx->mlp_w3q, xn, tn, wt, n0, batch,
                           ctx->mlp_gpart, ctx->mlp_lpart, ctx->stream);
    launch_mlp_adam(ctx->mlp_theta, ctx->mlp_m, ctx->mlp_v, ctx->mlp_gpart, nblk, (int)t, lr_t,
                    ctx->mlp_w1q, ctx->mlp_w3q, ctx->mlp_wa, nullptr, ctx->stream);
  }
  rc = check_launch();
  ctx->mlp_ready = true;
  if (!rc && train_loss) {
    std::vector<float> lp(nblk);
    if (hipMemcpyAsync(lp.data(), ctx->mlp_lpart, sizeof(float) * nblk, hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      rc = PHLX_E_HIP;
    } else {
      double s = 0.0;
      for (float v : lp) s += v;
      *train_loss = s / (double)batch;
    }
  }
  if (!rc && (val_rms_logT || val_rms_loge)) {
    // held-out tail: RMS error of log10 T and of the implied log10 E
    // (|d ln E / d ln T| times the log10 T error)
    float *y = xn;  // reuse: the normalised inputs of the training part are done
    launch_mlp_forward(ctx->g.mlp, nval, raw + 4 * ntr, y, nullptr, ctx->stream);
    std::vector<float> hy(nval), ht(nval), hw(nval);
    if (check_launch() ||
        hipMemcpyAsync(hy.data(), y, sizeof(float) * nval, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(ht.data(), tn + ntr, sizeof(float) * nval, hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(hw.data(), wt + ntr, sizeof(float) * nval, hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      rc = PHLX_E_HIP;
    } else {
      double s = 0.0, se = 0.0;
      for (long long i = 0; i < nval; i++) {
        const double d = ((double)hy[i] - (double)ht[i]) * norm[9];
        s += d * d;
        se += (double)hw[i] * d * d;
      }
      if (val_rms_logT) *val_rms_logT = std::sqrt(s / (double)nval);
      if (val_rms_loge) *val_rms_loge = std::sqrt(se / (double)nval);
    }
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(data);
  return rc;
}

extern "C" int phlx_helm_rhoe_given_guess_batch(phlx_ctx *ctx, int64_t n, const double *rho,
                                                const double *e, const double *abar,
                                                const double *zbar, double *T, double *P,
                                                double *sound, int32_t return_sound,
                                                int32_t guess_mode, int32_t *iters,
                                                int32_t *fallback) {
  if (!ctx || n < 0) return PHLX_E_ARG;
  if (guess_mode < GUESS_WARM || guess_mode > GUESS_BRACKET) return PHLX_E_ARG;
  if (!ctx->helm_loaded) return PHLX_E_TABLE;
  if (guess_mode == GUESS_MLP && !ctx->mlp_ready) return PHLX_E_STATE;
  if (n == 0) return PHLX_OK;
  if (!rho || !e || !abar || !zbar || !T) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  launch_helm_rhoe_guess(ctx->helm, ctx->g.mlp, guess_mode, n, rho, e, abar, zbar, T, P, sound,
                         return_sound, iters, fallback, ctx->stream);
  return check_launch();
}

extern "C" int phlx_set_eos_guess(phlx_ctx *ctx, int32_t mode, double resid_tol,
                                  int32_t count_iterations) {
  if (!ctx) return PHLX_E_ARG;
  if (mode < GUESS_WARM || mode > GUESS_PRED || !(resid_tol > 0.0)) return PHLX_E_ARG;
  if (ctx->cfg.eos == PHLX_EOS_IDEAL && !ctx->helm_loaded) return PHLX_E_CONFIG;
  // the surrogate is trained on Helmholtz states (abar, zbar inputs)
  if (mode == GUESS_MLP && (ctx->cfg.eos == PHLX_EOS_PIG || ctx->helm.pig)) return PHLX_E_CONFIG;
  if (mode == GUESS_MLP && !ctx->mlp_ready) return PHLX_E_STATE;
  // the guess variant of the update needs every block full of interior cells
  if (ctx->cfg.eos != PHLX_EOS_IDEAL && needs_guess_variant(mode, count_iterations != 0) &&
      (ctx->g.n1 % 64 || ctx->g.n2 % 4))
    return PHLX_E_CONFIG;
  if (mode == GUESS_MLP && ctx->g.eos_guess != GUESS_MLP) {
    // no previous network output yet: all-ones bytes are a NaN pattern
    HIPCHK(hipMemsetAsync(ctx->g.mlp_y, 0xFF, sizeof(float) * ctx->g.N, ctx->stream));
  }
  if (mode == GUESS_PRED && ctx->g.eos_guess != GUESS_PRED) {
    int rc = forget_predictor(ctx);
    if (rc) return rc;
  }
  ctx->g.eos_guess = mode;
  ctx->g.mlp.resid_tol = (float)resid_tol;
  ctx->eos_stats = count_iterations != 0;
  ctx->g.eos_count = ctx->eos_stats ? 1 : 0;
  return PHLX_OK;
}

extern "C" int phlx_set_eos_lin_tol(phlx_ctx *ctx, double lin_tol) {
  if (!ctx || !(lin_tol >= 0.0) || lin_tol > 1e-2) return PHLX_E_ARG;
  ctx->g.eos_lin_tol = lin_tol;
  return PHLX_OK;
}

// ------------------------------------------------------- thermal diffusion

// THERMAL_DIFFUSION_STS (source/source.F90:19804-21068).  The ideal gas needs
// a constant mean molecular weight (nas 0); Helmholtz diffuses with its
// carried T.  fix_t: six wall temperatures (NaN: zero-gradient ghost), only
// on non-periodic sides.
extern "C" int phlx_set_thermal_diffusion(phlx_ctx *ctx, int32_t enable, int32_t evaluate_dt,
                                          const double *fix_t) {
  if (!ctx || (enable != 0 && enable != 1) || (evaluate_dt != 0 && evaluate_dt != 1))
    return PHLX_E_ARG;
  Grid &g = ctx->g;
  if (!enable) {
    ctx->diff_on = false;
    return PHLX_OK;
  }
  const bool helm = ctx->cfg.eos != PHLX_EOS_IDEAL;
  if (!helm && g.nas != 0) return PHLX_E_CONFIG;
  double ft[6];
  for (int i = 0; i < 6; i++) {
    ft[i] = fix_t ? fix_t[i] : std::nan("");
    if (ft[i] == ft[i]) {
      const int d = i / 2;
      if (d >= g.sd || g.bc[i] == PHLX_BC_PERIODIC || !(ft[i] > 0.0) || !std::isfinite(ft[i]))
        return PHLX_E_ARG;
    }
  }
  HIPCHK(hipSetDevice(ctx->device));
  if (!g.e0) {
    const long long k3 = (g.sd == 3) ? g.n3 + 2 : 1;
    const size_t nk = (size_t)(g.n1 + 2) * (g.n2 + 2) * k3;
    int rc = dalloc(ctx, &g.kappa, nk);
    if (!rc) rc = dalloc(ctx, &g.e0, (size_t)g.N);
    if (!rc) rc = dalloc(ctx, &g.ejm2, (size_t)g.N);
    if (!rc) rc = dalloc(ctx, &g.me0, (size_t)g.N);
    if (!rc && !helm) rc = dalloc(ctx, &g.tdiff, (size_t)g.N);
    if (rc) return rc;
    if (helm) g.tdiff = g.temp;
  }
  for (int i = 0; i < 6; i++) g.fix_t[i] = ft[i];
  ctx->diff_eval = evaluate_dt != 0;
  ctx->diff_on = true;
  return PHLX_OK;
}

// kappa on cells -1 .. n of each axis (one ghost layer, x3 only in 3D),
// x1 fastest, as the application fills lgrid%kappa.
extern "C" int phlx_set_kappa(phlx_ctx *ctx, const double *kappa) {
  if (!ctx || !kappa) return PHLX_E_ARG;
  Grid &g = ctx->g;
  if (!g.kappa) return PHLX_E_STATE;
  const long long k3 = (g.sd == 3) ? g.n3 + 2 : 1;
  const size_t nk = (size_t)(g.n1 + 2) * (g.n2 + 2) * k3;
  for (size_t i = 0; i < nk; i++)
    if (!(kappa[i] > 0.0) || !std::isfinite(kappa[i])) return PHLX_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpyAsync(g.kappa, kappa, nk * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->kappa_set = true;
  return PHLX_OK;
}

static int parabolic_dt(phlx_ctx *ctx) {
  if (!ctx->kappa_set) return PHLX_E_STATE;
  if (const int m = materialize(ctx)) return m;
  const Grid &g = ctx->g;
  const double i1 = g.inv_dx[0], i2 = g.inv_dx[1], i3 = (g.sd == 3) ? g.inv_dx[2] : 0.0;
  ProfRec r;
  int on = kbegin(ctx, KID_DIFF, r);
  launch_parabolic(g, i1 * i1 + i2 * i2 + i3 * i3, ctx->stream);
  launch_reduce_to(g, ctx->scratch_val, ctx->stream);
  kend(ctx, on, r);
  int rc = check_launch();
  if (rc) return rc;
  double m = 0.0;
  HIPCHK(hipMemcpyAsync(&m, ctx->scratch_val, sizeof(m), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->dt_par = 1.0 / m;
  return PHLX_OK;
}

extern "C" int phlx_parabolic_dt(phlx_ctx *ctx, double *dt_par) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->diff_on) return PHLX_E_CONFIG;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = parabolic_dt(ctx);
  if (rc) return rc;
  if (dt_par) *dt_par = ctx->dt_par;
  return PHLX_OK;
}

// thermal_diffusion_step over dt/2 (source/source.F90:19949-20269): the
// stage count and the RKL2 coefficients on the host from the step's dt.
static int diffusion_step(phlx_ctx *ctx, int strang) {
  (void)strang;
  if (const int m = materialize(ctx)) return m;
  if (!ctx->kappa_set || !(ctx->dt_par > 0.0)) return PHLX_E_STATE;
  Scalars sc;
  HIPCHK(hipMemcpyAsync(&sc, ctx->g.sc, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (!sc.active) return PHLX_OK;
  const Grid &g = ctx->g;
  const bool helm = ctx->cfg.eos != PHLX_EOS_IDEAL;
  const double fac = sc.dt / ctx->dt_par, half_dt = 0.5 * sc.dt;
  const double sr = std::floor(0.5 * (std::sqrt(9.0 + 16.0 * fac) - 1.0));
  if (!(sr < 1e6)) return PHLX_E_ARG;  // dt / dt_par beyond any sensible stage count
  int s = 1 + (int)sr;
  if (s < 3) s = 3;
  const double w1 = 4.0 / ((double)(s * s + s) - 2.0);
  const double mut1 = (1.0 / 3.0) * w1;
  ProfRec r;
  int on = kbegin(ctx, KID_DIFF, r);
  if (!helm && !ctx->tdiff_valid) launch_diff_t_from_p(g, ctx->stream);
  launch_diff_stage(g, true, 0.0, 0.0, 0.0, half_dt * mut1, 0.0, ctx->stream);
  double bjm2 = 1.0 / 3.0, bjm1 = 1.0 / 3.0;
  for (int j = 2; j <= s; j++) {
    const double bj = ((double)(j * j + j) - 2.0) / (2.0 * j * (j + 1.0));
    const double muj = ((2.0 * j - 1.0) / j) * bj / bjm1;
    const double mutj = muj * w1;
    const double nuj = -(((j - 1.0) / j) * bj / bjm2);
    const double gammatj = -(1.0 - bjm1) * mutj;
    launch_diff_temp(g, helm, ctx->stream);
    launch_diff_stage(g, false, muj, nuj, 1.0 - muj - nuj, half_dt * mutj, half_dt * gammatj,
                      ctx->stream);
    bjm2 = bjm1;
    bjm1 = bj;
  }
  launch_diff_final(g, helm, ctx->stream);
  if (ctx->kt.pack) {
    // the primitives moved past U
    ctx->kt.pack(g, ctx->stream);
    launch_flags(g.sc, ~SC_U_CUR, 0, ctx->stream);
  }
  kend(ctx, on, r);
  int rc = helm ? forget_predictor(ctx) : PHLX_OK;
  if (rc) return rc;
  rc = check_launch();
  if (rc) return rc;
  ctx->tdiff_valid = true;
  ctx->diff_stages = s;
  return PHLX_OK;
}

extern "C" int phlx_diffusion_step(phlx_ctx *ctx, int32_t strang) {
  if (!ctx || (strang != 1 && strang != 2)) return PHLX_E_ARG;
  if (!ctx->initialized) return PHLX_E_STATE;
  if (!ctx->diff_on) return PHLX_E_CONFIG;
  HIPCHK(hipSetDevice(ctx->device));
  return diffusion_step(ctx, strang);
}

extern "C" int phlx_diffusion_stages(phlx_ctx *ctx, int32_t *stages) {
  if (!ctx || !stages) return PHLX_E_ARG;
  *stages = ctx->diff_stages;
  return PHLX_OK;
}

// USE_SHOCK_FLATTENING with eps_sf_make (source/source.F90:8073-8131): the
// detector needs three ghost cells, so every axis holds at least 3 cells;
// FP64 only.  The flag arrays are made on the first enable.
extern "C" int phlx_set_shock_flattening(phlx_ctx *ctx, int32_t enable, double eps_sf) {
  if (!ctx || (enable != 0 && enable != 1)) return PHLX_E_ARG;
  Grid &g = ctx->g;
  if (!enable) {
    g.sflat = 0;
    return PHLX_OK;
  }
  if (!(eps_sf > 0.0) || !std::isfinite(eps_sf)) return PHLX_E_ARG;
  if (ctx->cfg.precision != PHLX_PREC_FP64) return PHLX_E_CONFIG;
  if (g.n1 < 3 || g.n2 < 3 || (g.sd == 3 && g.n3 < 3)) return PHLX_E_CONFIG;
  HIPCHK(hipSetDevice(ctx->device));
  if (!g.sf_flag) {
    const long long e3 = (g.sd == 3) ? g.n3 + 4 : 1, f3 = (g.sd == 3) ? g.n3 + 2 : 1;
    const size_t nd = (size_t)(g.n1 + 4) * (g.n2 + 4) * e3;
    const size_t nfl = (size_t)(g.n1 + 2) * (g.n2 + 2) * f3;
    unsigned char *det = nullptr, *flag = nullptr;
    int rc = dalloc(ctx, &det, nd);
    if (rc) return rc;
    rc = dalloc(ctx, &flag, nfl);
    if (rc) return rc;
    g.sf_det = det;
    g.sf_flag = flag;
  }
  g.eps_sf = eps_sf;
  g.sflat = 1;
  return PHLX_OK;
}

extern "C" int phlx_eos_stats(phlx_ctx *ctx, int64_t *newton_iters, int64_t *solves,
                              int64_t *fallbacks, int32_t reset) {
  if (!ctx) return PHLX_E_ARG;
  if (!ctx->g.eos_part) return PHLX_E_CONFIG;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<unsigned int> part(3 * (size_t)ctx->g.nred);
  HIPCHK(hipMemcpyAsync(part.data(), ctx->g.eos_part, sizeof(unsigned int) * part.size(),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  long long s[3] = {0, 0, 0};
  for (size_t b = 0; b < part.size(); b += 3)
    for (int q = 0; q < 3; q++) s[q] += part[b + q];
  if (newton_iters) *newton_iters = s[0];
  if (solves) *solves = s[1];
  if (fallbacks) *fallbacks = s[2];
  if (reset) {
    HIPCHK(hipMemsetAsync(ctx->g.eos_part, 0, sizeof(unsigned int) * part.size(), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return PHLX_OK;
}
