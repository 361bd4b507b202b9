/*
 * phlx_oracle.c -- CPU FP64 restatement of Phlegethon's hydro_step path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness checker for the HIP product and
 * the CPU baseline timed beside it.  Nothing in the product links or calls
 * this file.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Every routine restates the reference algorithm from source/source.F90 and
 * keeps its floating-point operation order (Fortran evaluates a*b*c as
 * (a*b)*c and a+b+c as (a+b)+c; no FMA contraction), so that a faithful
 * device implementation agrees to round-off.  Citations are file:line into
 * the reference tree.
 *
 * Parity pins (tests/test_oracle_*.py): Riemann consistency F(q,q) against
 * the closed-form physical flux, Sod / Brio-Wu style shock tubes against the
 * exact solution, hydrostatic rest state preserved bitwise by the
 * well-balanced scheme, rotation (x1<->x2) invariance of a full 2D run,
 * div B at round-off under CT, second-order convergence of smooth sound
 * and circularly polarised Alfven waves, reconstruction exactness.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CONST_RGAS 8.31446261815324e7 /* source.F90:605 */
/* HELMHOLTZ_EOS or PIG_EOS: a tabulated EoS with a carried T (the table
 * object says which interpolation applies) */
#define TABULATED(g) ((g)->c.eos == ORC_EOS_HELMHOLTZ || (g)->c.eos == ORC_EOS_PIG)

struct orc_grid {
  orc_config c;
  int nv, ng, sd, mhd;
  int n[3];        /* interior cells per direction (n[2] = 1 in 2D) */
  int e1, e2, e3;  /* ghosted extents */
  int g3;          /* ghost depth along x3 (0 in 2D) */
  int iv[3], ip, ias1;
  double dx[3], inv_dx[3];
  double gm, gmm1, igmm1, inv_mu;
  double *prim, *bcc, *temp, *gamf; /* ghosted, AoS (nv / 3 / 1 / 2 per cell) */
  unsigned char *flat;               /* ghosted: USE_SHOCK_FLATTENING cell flags */
  double *eint;                     /* interior */
  double *state, *state0, *res;     /* interior AoS [cell][nv] */
  double *bf[3], *b0f[3];           /* face B, interior faces */
  double *emf[3], *sgn[3];          /* face EMFs [face][3] and contact signs, 1 ghost */
  int fe[3][3];                     /* extents of the ghosted face arrays */
  double *cor[3];                   /* corner EMFs along x1, x2, x3 */
  double *eq_rho_cc, *eq_p_cc, *eq_rho_f, *eq_p_f, *grav, *phi_cc, *phi_f, *edot;
  double *eq_ge_cc, *eq_ge_f;       /* FASTEOS WB: equilibrium gamma_e */
  const orc_helm *helm;             /* Helmholtz table (not owned) */
  /* THERMAL_DIFFUSION_STS: kappa on cells -1..n (one ghost layer), RKL2
   * registers e_0, e_jm2, Me_0 and the residual Me (interior) */
  double *kappa, *e0, *ejm2, *me0, *me;
  double dt_par;
  double time, dt, smax;
  long step;
};

/* ---------------------------------------------------------------- indexing */

static inline long cidx(const orc_grid *g, int i, int j, int k) {
  return ((long)(k + g->g3) * g->e2 + (j + g->ng)) * g->e1 + (i + g->ng);
}
static inline long iidx(const orc_grid *g, int i, int j, int k) {
  return ((long)k * g->n[1] + j) * g->n[0] + i;
}
/* interior face arrays: direction d has n[d]+1 faces along d */
static inline long fidx(const orc_grid *g, int d, int i, int j, int k) {
  int f1 = g->n[0] + (d == 0), f2 = g->n[1] + (d == 1);
  return ((long)k * f2 + j) * f1 + i;
}
/* ghosted face arrays (1 layer; along x3 in 2D no ghost) */
static inline long gidx(const orc_grid *g, int d, int i, int j, int k) {
  int lo3 = (g->sd == 3) ? -1 : 0;
  return ((long)(k - lo3) * g->fe[d][1] + (j + 1)) * g->fe[d][0] + (i + 1);
}
/* corner arrays: cor[2] (x3-edges) n1+1 x n2+1 x n3; cor[0] n1 x n2+1 x n3+1;
 * cor[1] n1+1 x n2 x n3+1 */
static inline long kidx(const orc_grid *g, int d, int i, int j, int k) {
  int c1 = g->n[0] + (d != 0), c2 = g->n[1] + (d != 1);
  return ((long)k * c2 + j) * c1 + i;
}

/* ------------------------------------------------------------- lifecycle */

orc_grid *orc_create(const orc_config *cfg) {
  orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
  if (!g) return NULL;
  g->c = *cfg;
  g->sd = cfg->sdims;
  g->nv = cfg->sdims + 2 + cfg->nas;            /* source.F90:472-477 */
  /* LIM5TH reads i +- 2 of cells lx1-1..ux1+1; the shock detector reads
   * i +- 1 of cells lx1-2..ux1+2 (with the reference's ngc = 2 its outer
   * ring reads past the ghosts; here the ghosts are filled one deeper) */
  g->ng = (cfg->recon == ORC_LIM5TH || cfg->shock_flattening) ? 3 : 2;
  g->mhd = (cfg->flux == ORC_HLLD || cfg->flux == ORC_LHLLD);
  g->n[0] = cfg->nx1;
  g->n[1] = cfg->nx2;
  g->n[2] = (cfg->sdims == 3) ? cfg->nx3 : 1;
  g->g3 = (cfg->sdims == 3) ? g->ng : 0;
  g->e1 = g->n[0] + 2 * g->ng;
  g->e2 = g->n[1] + 2 * g->ng;
  g->e3 = g->n[2] + 2 * g->g3;
  g->iv[0] = 1;
  g->iv[1] = 2;
  g->iv[2] = (cfg->sdims == 3) ? 3 : 2;          /* i_vx3 = i_vx2 in 2D, :487-491 */
  g->ip = g->iv[2] + 1;
  g->ias1 = 2 + cfg->sdims;
  /* source.F90:1725-1734 */
  g->dx[0] = (cfg->x1u - cfg->x1l) / (double)cfg->nx1;
  g->dx[1] = (cfg->x2u - cfg->x2l) / (double)cfg->nx2;
  g->dx[2] = (cfg->sdims == 3) ? (cfg->x3u - cfg->x3l) / (double)cfg->nx3 : 1.0;
  for (int d = 0; d < 3; d++) g->inv_dx[d] = 1.0 / g->dx[d];
  g->gm = cfg->gamma;
  g->gmm1 = g->gm - 1.0;
  g->igmm1 = 1.0 / g->gmm1;
  g->inv_mu = 1.0 / cfg->mu;

  long nc = (long)g->e1 * g->e2 * g->e3;
  long ni = (long)g->n[0] * g->n[1] * g->n[2];
  g->prim = (double *)calloc(nc * g->nv, sizeof(double));
  g->bcc = (double *)calloc(nc * 3, sizeof(double));
  g->temp = (double *)calloc(nc, sizeof(double));
  g->gamf = (double *)calloc(nc * 2, sizeof(double));
  g->flat = (unsigned char *)calloc(nc, 1);
  g->eint = (double *)calloc(ni, sizeof(double));
  g->state = (double *)calloc(ni * g->nv, sizeof(double));
  g->state0 = (double *)calloc(ni * g->nv, sizeof(double));
  g->res = (double *)calloc(ni * g->nv, sizeof(double));
  for (int d = 0; d < 3; d++) {
    long nf = (long)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
    g->bf[d] = (double *)calloc(nf, sizeof(double));
    g->b0f[d] = (double *)calloc(nf, sizeof(double));
    for (int a = 0; a < 3; a++) g->fe[d][a] = g->n[a] + 2 + (a == d);
    if (g->sd == 2) g->fe[d][2] = 1 + (d == 2);
    long ng_ = (long)g->fe[d][0] * g->fe[d][1] * g->fe[d][2];
    g->emf[d] = (double *)calloc(ng_ * 3, sizeof(double));
    g->sgn[d] = (double *)calloc(ng_, sizeof(double));
    long nk = (long)(g->n[0] + (d != 0)) * (g->n[1] + (d != 1)) * (g->n[2] + (d != 2));
    g->cor[d] = (double *)calloc(nk, sizeof(double));
  }
  int ny = g->n[1];
  g->eq_rho_cc = (double *)calloc(ny, sizeof(double));
  g->eq_p_cc = (double *)calloc(ny, sizeof(double));
  g->eq_rho_f = (double *)calloc(ny + 1, sizeof(double));
  g->eq_p_f = (double *)calloc(ny + 1, sizeof(double));
  g->grav = (double *)calloc(ny, sizeof(double));
  g->phi_cc = (double *)calloc(ny, sizeof(double));
  g->phi_f = (double *)calloc(ny + 1, sizeof(double));
  g->edot = (double *)calloc(ny, sizeof(double));
  g->eq_ge_cc = (double *)calloc(ny, sizeof(double));
  g->eq_ge_f = (double *)calloc(ny + 1, sizeof(double));
#ifdef _OPENMP
  if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
  return g;
}

void orc_destroy(orc_grid *g) {
  if (!g) return;
  free(g->prim); free(g->bcc); free(g->temp); free(g->gamf); free(g->eint); free(g->flat);
  free(g->kappa); free(g->e0); free(g->ejm2); free(g->me0); free(g->me);
  free(g->state); free(g->state0); free(g->res);
  for (int d = 0; d < 3; d++) {
    free(g->bf[d]); free(g->b0f[d]); free(g->emf[d]); free(g->sgn[d]); free(g->cor[d]);
  }
  free(g->eq_rho_cc); free(g->eq_p_cc); free(g->eq_rho_f); free(g->eq_p_f);
  free(g->grav); free(g->phi_cc); free(g->phi_f); free(g->edot);
  free(g->eq_ge_cc); free(g->eq_ge_f);
  free(g);
}

int orc_nvars(const orc_grid *g) { return g->nv; }

int orc_set_state(orc_grid *g, const double *prim, const double *eint,
                  const double *temp, const double *b1, const double *b2,
                  const double *b3) {
  long ni = (long)g->n[0] * g->n[1] * g->n[2];
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        if (prim)
          for (int v = 0; v < g->nv; v++) g->prim[c * g->nv + v] = prim[v * ni + q];
        if (eint) g->eint[q] = eint[q];
        if (temp) g->temp[c] = temp[q];
      }
  const double *bs[3] = {b1, b2, b3};
  for (int d = 0; d < g->sd; d++) {
    if (!bs[d]) continue;
    long nf = (long)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
    memcpy(g->bf[d], bs[d], nf * sizeof(double));
  }
  return 0;
}

int orc_get_state(const orc_grid *g, double *prim, double *eint, double *temp,
                  double *b1, double *b2, double *b3, double *bcc) {
  long ni = (long)g->n[0] * g->n[1] * g->n[2];
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        if (prim)
          for (int v = 0; v < g->nv; v++) prim[v * ni + q] = g->prim[c * g->nv + v];
        if (eint) eint[q] = g->eint[q];
        if (temp) temp[q] = g->temp[c];
        if (bcc)
          for (int v = 0; v < 3; v++) bcc[v * ni + q] = g->bcc[c * 3 + v];
      }
  double *bs[3] = {b1, b2, b3};
  for (int d = 0; d < g->sd; d++) {
    if (!bs[d]) continue;
    long nf = (long)(g->n[0] + (d == 0)) * (g->n[1] + (d == 1)) * (g->n[2] + (d == 2));
    memcpy(bs[d], g->bf[d], nf * sizeof(double));
  }
  return 0;
}

int orc_set_profiles(orc_grid *g, const double *eq_rho_cc, const double *eq_p_cc,
                     const double *eq_rho_f, const double *eq_p_f,
                     const double *grav_cc, const double *phi_cc,
                     const double *phi_f, const double *edot_cc) {
  int ny = g->n[1];
  if (eq_rho_cc) memcpy(g->eq_rho_cc, eq_rho_cc, ny * sizeof(double));
  if (eq_p_cc) memcpy(g->eq_p_cc, eq_p_cc, ny * sizeof(double));
  if (eq_rho_f) memcpy(g->eq_rho_f, eq_rho_f, (ny + 1) * sizeof(double));
  if (eq_p_f) memcpy(g->eq_p_f, eq_p_f, (ny + 1) * sizeof(double));
  if (grav_cc) memcpy(g->grav, grav_cc, ny * sizeof(double));
  if (phi_cc) memcpy(g->phi_cc, phi_cc, ny * sizeof(double));
  if (phi_f) memcpy(g->phi_f, phi_f, (ny + 1) * sizeof(double));
  if (edot_cc) memcpy(g->edot, edot_cc, ny * sizeof(double));
  return 0;
}

int orc_set_eos_profiles(orc_grid *g, const double *eq_gammae_cc, const double *eq_gammae_f) {
  int ny = g->n[1];
  if (eq_gammae_cc) memcpy(g->eq_ge_cc, eq_gammae_cc, ny * sizeof(double));
  if (eq_gammae_f) memcpy(g->eq_ge_f, eq_gammae_f, (ny + 1) * sizeof(double));
  return 0;
}

int orc_set_helm(orc_grid *g, const orc_helm *h) {
  g->helm = h;
  return 0;
}

/* gamma_e, gamma_c of the interior cells, [2][cell] */
int orc_get_gammaf(const orc_grid *g, double *out) {
  long ni = (long)g->n[0] * g->n[1] * g->n[2];
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        out[q] = g->gamf[c * 2];
        out[ni + q] = g->gamf[c * 2 + 1];
      }
  return 0;
}

double orc_get_scalar(const orc_grid *g, int which) {
  switch (which) {
    case 0: return g->time;
    case 1: return g->dt;
    case 2: return g->smax;
    case 3: return (double)g->step;
  }
  return 0.0;
}

void orc_set_scalar(orc_grid *g, int which, double v) {
  switch (which) {
    case 0: g->time = v; break;
    case 1: g->dt = v; break;
    case 2: g->smax = v; break;
    case 3: g->step = (long)v; break;
  }
}

/* ------------------------------------------------------- ghost-zone fill */

/* communicate_ndarray for a single periodic brick (source.F90:2361-2446):
 * direction by direction, the other directions' ranges extended by the ghost
 * depth when corners are requested. */
static void halo_periodic(orc_grid *g, double *a, int nc) {
  int ng = g->ng;
  for (int d = 0; d < g->sd; d++) {
    if (g->c.bc[2 * d] != ORC_BC_PERIODIC) continue;
    int lo[3], hi[3];
    for (int a2 = 0; a2 < 3; a2++) {
      lo[a2] = (a2 < g->sd) ? -ng : 0;
      hi[a2] = (a2 < g->sd) ? g->n[a2] + ng : g->n[a2];
    }
    int n = g->n[d];
    for (int edge = 0; edge < 2; edge++) {
      int s0 = edge == 0 ? n - ng : 0;
      int r0 = edge == 0 ? -ng : n;
      int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
      l2[d] = s0;
      h2[d] = s0 + ng;
      for (int k = l2[2]; k < h2[2]; k++)
        for (int j = l2[1]; j < h2[1]; j++)
          for (int i = l2[0]; i < h2[0]; i++) {
            int ii[3] = {i, j, k};
            ii[d] = ii[d] - s0 + r0;
            long src = cidx(g, i, j, k), dst = cidx(g, ii[0], ii[1], ii[2]);
            for (int v = 0; v < nc; v++) a[dst * nc + v] = a[src * nc + v];
          }
    }
  }
}

/* Cartesian physical boundaries (source.F90:7532-7956): ghost = fac * mirror */
static void physical_bc(orc_grid *g) {
  int ng = g->ng;
  for (int d = 0; d < g->sd; d++) {
    for (int side = 0; side < 2; side++) {
      int bc = g->c.bc[2 * d + side];
      if (bc == ORC_BC_PERIODIC) continue;
      double fac[16], facb[3];
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = (a < g->sd) ? -ng : 0;
        hi[a] = (a < g->sd) ? g->n[a] + ng : g->n[a];
      }
      int n = g->n[d];
      /* iterate over the transverse plane (full ghosted ranges) */
      int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
      for (int b = lo[t2]; b < hi[t2]; b++)
        for (int a = lo[t1]; a < hi[t1]; a++) {
          for (int v = 0; v < g->nv; v++) fac[v] = 1.0;
          for (int v = 0; v < 3; v++) facb[v] = 1.0;
          if (bc == ORC_BC_REFLECTIVE) {
            fac[g->iv[d]] = -1.0;
            facb[d] = -1.0;
          } else if (bc == ORC_BC_DIODE) {
            int ii[3];
            ii[t1] = a; ii[t2] = b; ii[d] = side == 0 ? 0 : n - 1;
            double vn = g->prim[cidx(g, ii[0], ii[1], ii[2]) * g->nv + g->iv[d]];
            if (side == 0) vn = -vn;
            double f = (vn < 0.0) ? -1.0 : 1.0;
            fac[g->iv[d]] = f;
            facb[d] = f;
          }
          for (int gg = 0; gg < ng; gg++) {
            int gi = side == 0 ? -1 - gg : n + gg;
            int ri = side == 0 ? gg : n - 1 - gg;
            int id[3], ir[3];
            id[t1] = a; id[t2] = b; id[d] = gi;
            ir[t1] = a; ir[t2] = b; ir[d] = ri;
            long cd = cidx(g, id[0], id[1], id[2]), cr = cidx(g, ir[0], ir[1], ir[2]);
            for (int v = 0; v < g->nv; v++) g->prim[cd * g->nv + v] = fac[v] * g->prim[cr * g->nv + v];
            if (g->mhd)
              for (int v = 0; v < g->sd; v++) g->bcc[cd * 3 + v] = facb[v] * g->bcc[cr * 3 + v];
            if (TABULATED(g)) g->temp[cd] = g->temp[cr];
            if (g->c.fasteos) {
              g->gamf[cd * 2] = g->gamf[cr * 2];
              g->gamf[cd * 2 + 1] = g->gamf[cr * 2 + 1];
            }
          }
        }
    }
  }
}

/* EMF / contact-sign halo: communicate_ndarray(...,gst=1,corners=.false.)
 * on the face-centred arrays (source.F90:11370-11394) plus the EMF boundary
 * conditions (source.F90:11396-11964, reflective and outflow). */
static void emf_halo(orc_grid *g) {
  for (int f = 0; f < g->sd; f++) {        /* face array of direction f */
    double *E = g->emf[f], *S = g->sgn[f];
    int up[3];
    for (int a = 0; a < 3; a++) up[a] = g->n[a] - 1 + (a == f);
    /* periodic copies: bounds i1=0 .. i2=up[d] in the communicated dim */
    for (int d = 0; d < g->sd; d++) {
      if (g->c.bc[2 * d] != ORC_BC_PERIODIC) continue;
      int lo[3] = {0, 0, 0}, hi[3] = {up[0] + 1, up[1] + 1, up[2] + 1};
      for (int edge = 0; edge < 2; edge++) {
        int s0 = edge == 0 ? up[d] : 0;
        int r0 = edge == 0 ? -1 : up[d] + 1;
        int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
        l2[d] = s0;
        h2[d] = s0 + 1;
        for (int k = l2[2]; k < h2[2]; k++)
          for (int j = l2[1]; j < h2[1]; j++)
            for (int i = l2[0]; i < h2[0]; i++) {
              int ii[3] = {i, j, k};
              ii[d] = r0;
              long s = gidx(g, f, i, j, k), t = gidx(g, f, ii[0], ii[1], ii[2]);
              for (int v = 0; v < 3; v++) E[t * 3 + v] = E[s * 3 + v];
              S[t] = S[s];
            }
      }
    }
  }
  /* physical EMF boundaries: for boundary direction d, the face arrays of
   * the two other directions get one ghost layer along d. */
  for (int d = 0; d < g->sd; d++)
    for (int side = 0; side < 2; side++) {
      int bc = g->c.bc[2 * d + side];
      if (bc == ORC_BC_PERIODIC) continue;
      double facb[3] = {1.0, 1.0, 1.0};
      if (bc == ORC_BC_REFLECTIVE) {
        facb[0] = facb[1] = facb[2] = -1.0;
        facb[d] = 1.0;
      }
      for (int f = 0; f < g->sd; f++) {
        if (f == d) continue;
        int hi[3];
        for (int a = 0; a < 3; a++) hi[a] = g->n[a] + (a == f);
        int gi = side == 0 ? -1 : g->n[d];
        int ri = side == 0 ? 0 : g->n[d] - 1;
        int l2[3] = {0, 0, 0}, h2[3] = {hi[0], hi[1], hi[2]};
        l2[d] = ri;
        h2[d] = ri + 1;
        for (int k = l2[2]; k < h2[2]; k++)
          for (int j = l2[1]; j < h2[1]; j++)
            for (int i = l2[0]; i < h2[0]; i++) {
              int ii[3] = {i, j, k};
              ii[d] = gi;
              long s = gidx(g, f, i, j, k), t = gidx(g, f, ii[0], ii[1], ii[2]);
              if (bc == ORC_BC_REFLECTIVE || bc == ORC_BC_OUTFLOW)
                for (int v = 0; v < 3; v++) g->emf[f][t * 3 + v] = facb[v] * g->emf[f][s * 3 + v];
              g->sgn[f][t] = g->sgn[f][s];
            }
      }
    }
}

/* --------------------------------------------------------- reconstruction */

/* LIM2ND_REC (source.F90:8141-8181): harmonic-mean limited slope */
static inline void rec_lim2nd(double qm, double qc, double qp, double *l, double *r) {
  double slopef = qp - qc, slopeb = qc - qm;
  double prod = slopef * slopeb, slope = 0.0;
  if (prod > 0.0) slope = prod / (slopef + slopeb);
  double tmp = 0.5 * slope;
  *l = qc - tmp;
  *r = qc + tmp;
}

/* PPH_REC unlimited parabola (source.F90:8183-8193) */
static inline void rec_pph(double qm, double qc, double qp, double *l, double *r) {
  double tmp = (5.0 / 6.0) * qc;
  *l = (1.0 / 3.0) * qm + tmp - (1.0 / 6.0) * qp;
  *r = -(1.0 / 6.0) * qm + tmp + (1.0 / 3.0) * qp;
}

/* PPH_REC for advected scalars: clip + extremum limiter (source.F90:8194-8243) */
static inline void rec_pph_lim(double qm, double qc, double qp, double *l, double *r) {
  double tmp = (5.0 / 6.0) * qc;
  double am = (1.0 / 3.0) * qm + tmp - (1.0 / 6.0) * qp;
  double ap = -(1.0 / 6.0) * qm + tmp + (1.0 / 3.0) * qp;
  double mn = fmin(qm, qc), mx = fmax(qm, qc);
  if (am < mn) am = mn;
  if (am > mx) am = mx;
  mn = fmin(qc, qp);
  mx = fmax(qc, qp);
  if (ap < mn) ap = mn;
  if (ap > mx) ap = mx;
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  *l = am + qc;
  *r = ap + qc;
}

/* LIM5TH_REC: 5th-order face values clipped to the neighbour range + the
 * extremum limiter (source.F90:8261-8315, b :8401-8455, gamma :8607-8711) */
static inline void rec_lim5(double qmm, double qm, double qc, double qp, double qpp, double *l,
                            double *r) {
  const double o420 = 1.0 / 420.0;
  double tmp = 329.0 * qc;
  double am = (-21.0 * qmm + 189.0 * qm + tmp - 91.0 * qp + 14.0 * qpp) * o420;
  double ap = (14.0 * qmm - 91.0 * qm + tmp + 189.0 * qp - 21.0 * qpp) * o420;
  double mn = fmin(qm, qc), mx = fmax(qm, qc);
  if (am < mn) am = mn;
  if (am > mx) am = mx;
  mn = fmin(qc, qp);
  mx = fmax(qc, qp);
  if (ap < mn) ap = mn;
  if (ap > mx) ap = mx;
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  *l = am + qc;
  *r = ap + qc;
}

void orc_reconstruct(int scheme, int limited_scalar, int n, const double *q,
                     double *qm, double *qp) {
  if (scheme == ORC_LIM5TH) {
    for (int c = 2; c < n - 2; c++)
      rec_lim5(q[c - 2], q[c - 1], q[c], q[c + 1], q[c + 2], &qm[c], &qp[c]);
    return;
  }
  for (int c = 1; c < n - 1; c++) {
    if (scheme == ORC_LIM2ND)
      rec_lim2nd(q[c - 1], q[c], q[c + 1], &qm[c], &qp[c]);
    else if (limited_scalar)
      rec_pph_lim(q[c - 1], q[c], q[c + 1], &qm[c], &qp[c]);
    else
      rec_pph(q[c - 1], q[c], q[c + 1], &qm[c], &qp[c]);
  }
}

/* ------------------------------------------------------------ Riemann */

typedef struct {
  int sd, nv, nas, flux, all_mach, fasteos;
  double gm, igmm1, mach_cutoff;
  const orc_helm *helm; /* Helmholtz without FASTEOS: per-face T(rho, P) */
} rparams;

/* Helmholtz without FASTEOS: rho*e and the sound speed of one face state
 * from the Newton T(rho, P) started at the face guess Tg (HLLC
 * source.F90:15776-15780, :15891-15895; HLLD :14172-14176 and the others
 * alike); ye and abar from the state's advected scalars. */
static void face_helm(const rparams *P, const double *q, double Tg, double *rhoe, double *sound) {
  const int ia = 2 + P->sd, ie = P->sd + 1;
  const double rho = q[0], p = q[ie];
  /* PIG_EOS without advected scalars: the evaluation ignores Ye, abar */
  const double ye = (P->nv > ia + 1) ? q[ia] : 1.0;
  const double abar = (P->nv > ia + 1) ? 1.0 / q[ia + 1] : 1.0;
  const double zbar = ye * abar;
  double T = Tg, eint;
  orc_helm_rhoP_given(P->helm, rho, p, abar, zbar, &T, &eint, sound, 1);
  *rhoe = eint * rho;
}

static rparams make_rp(const orc_config *c) {
  rparams p;
  p.sd = c->sdims;
  p.nas = c->nas;
  p.nv = c->sdims + 2 + c->nas;
  p.flux = c->flux;
  p.all_mach = c->all_mach;
  p.fasteos = c->fasteos;
  p.gm = c->gamma;
  p.igmm1 = 1.0 / (c->gamma - 1.0);
  p.mach_cutoff = c->mach_cutoff;
  p.helm = NULL;
  return p;
}

/* HLLC / LHLLC, Cartesian (source.F90:15611-16260). q = (rho, v1, v2, [v3],
 * p, scalars); gL/gR = (gamma_e, gamma_c) face values for FASTEOS. */
static void riemann_hydro_face(const rparams *P, const double *qL, const double *qR,
                               const double *gL, const double *gR, double Tg, int hll,
                               double *flux) {
  const int sd = P->sd, nv = P->nv;
  const int im1 = 1, im2 = 2, im3 = (sd == 3) ? 3 : 2, ie = im3 + 1, ia = 2 + sd;
  double rhoL = qL[0], inv_rhoL = 1.0 / rhoL;
  double vx1L = qL[1], vx2L = qL[2], vx3L = (sd == 3) ? qL[3] : 0.0;
  double v2L = vx1L * vx1L + vx2L * vx2L + vx3L * vx3L;
  double pL = qL[ie], eint, cL, cR;
  if (P->fasteos) {
    eint = pL / (gL[0] - 1.0);
    cL = sqrt(gL[1] * pL * inv_rhoL);
  } else if (P->helm) {
    face_helm(P, qL, Tg, &eint, &cL);
  } else {
    eint = P->igmm1 * pL;
    cL = sqrt(P->gm * pL * inv_rhoL);
  }
  double rhoeL = eint + 0.5 * rhoL * v2L;
  double rhoR = qR[0], inv_rhoR = 1.0 / rhoR;
  double vx1R = qR[1], vx2R = qR[2], vx3R = (sd == 3) ? qR[3] : 0.0;
  double v2R = vx1R * vx1R + vx2R * vx2R + vx3R * vx3R;
  double pR = qR[ie];
  if (P->fasteos) {
    eint = pR / (gR[0] - 1.0);
    cR = sqrt(gR[1] * pR * inv_rhoR);
  } else if (P->helm) {
    face_helm(P, qR, Tg, &eint, &cR);
  } else {
    eint = P->igmm1 * pR;
    cR = sqrt(P->gm * pR * inv_rhoR);
  }
  double rhoeR = eint + 0.5 * rhoR * v2R;
  double max_c = fmax(cL, cR);
  double sL = fmin(vx1L, vx1R) - max_c;
  double sR = fmax(vx1L, vx1R) + max_c;
  double dsuL = sL - vx1L, dsuR = sR - vx1R;
  double ustar = (pR - pL + rhoL * vx1L * dsuL - rhoR * vx1R * dsuR) / (rhoL * dsuL - rhoR * dsuR);
  double rhostarL = rhoL * (dsuL / (sL - ustar));
  double rhoestarL = rhostarL * (rhoeL * inv_rhoL + (ustar - vx1L) * (ustar + pL * inv_rhoL / dsuL));
  double rhostarR = rhoR * (dsuR / (sR - ustar));
  double rhoestarR = rhostarR * (rhoeR * inv_rhoR + (ustar - vx1R) * (ustar + pR * inv_rhoR / dsuR));

  if (hll) {
    /* USE_SHOCK_FLATTENING: two-wave HLL flux on a flattened face
     * (source.F90:15929-16003) */
    double UL[16], UR[16], fL[16], fR[16];
    sL = fmin(0.0, sL);
    sR = fmax(0.0, sR);
    UL[0] = rhoL; UL[im1] = rhoL * vx1L; UL[im2] = rhoL * vx2L;
    if (sd == 3) UL[im3] = rhoL * vx3L;
    UL[ie] = rhoeL;
    for (int v = ia; v < nv; v++) UL[v] = rhoL * qL[v];
    UR[0] = rhoR; UR[im1] = rhoR * vx1R; UR[im2] = rhoR * vx2R;
    if (sd == 3) UR[im3] = rhoR * vx3R;
    UR[ie] = rhoeR;
    for (int v = ia; v < nv; v++) UR[v] = rhoR * qR[v];
    fL[0] = rhoL * vx1L; fL[im1] = rhoL * vx1L * vx1L + pL; fL[im2] = rhoL * vx1L * vx2L;
    if (sd == 3) fL[im3] = rhoL * vx1L * vx3L;
    fL[ie] = (rhoeL + pL) * vx1L;
    for (int v = ia; v < nv; v++) fL[v] = UL[v] * vx1L;
    fR[0] = rhoR * vx1R; fR[im1] = rhoR * vx1R * vx1R + pR; fR[im2] = rhoR * vx1R * vx2R;
    if (sd == 3) fR[im3] = rhoR * vx1R * vx3R;
    fR[ie] = (rhoeR + pR) * vx1R;
    for (int v = ia; v < nv; v++) fR[v] = UR[v] * vx1R;
    double dummy = 1.0 / (sR - sL), phi = sL * sR;
    for (int v = 0; v < nv; v++) flux[v] = dummy * (sR * fL[v] - sL * fR[v] + phi * (UR[v] - UL[v]));
    return;
  }

  if (P->flux == ORC_LHLLC) {
    /* Minoshima & Miyoshi low-dissipation pressure (source.F90:16014-16030) */
    double machL = sqrt(v2L) / cL, machR = sqrt(v2R) / cR;
    double chi = fmin(1.0, fmax(machL, machR));
    chi = fmax(P->mach_cutoff, chi);
    double phi = chi * (2.0 - chi);
    double rhobar = 0.5 * (rhoL + rhoR), cbar = 0.5 * (cL + cR);
    double pstar = 0.5 * (pL + pR) - 0.5 * phi * rhobar * cbar * (vx1R - vx1L);
    double sgn = copysign(1.0, ustar);
    double hp = 0.5 * (1.0 + sgn), hm = 0.5 * (1.0 - sgn);
    double rhostar = hp * rhostarL + hm * rhostarR;
    if (P->all_mach && sL > 0.0) {
      flux[0] = rhoL * vx1L;
      flux[im1] = rhoL * vx1L * vx1L + pL;
      flux[im2] = rhoL * vx1L * vx2L;
      if (sd == 3) flux[im3] = rhoL * vx1L * vx3L;
      flux[ie] = (rhoeL + pL) * vx1L;
      for (int v = ia; v < nv; v++) flux[v] = rhoL * qL[v] * vx1L;
    } else if (P->all_mach && sR < 0.0) {
      flux[0] = rhoR * vx1R;
      flux[im1] = rhoR * vx1R * vx1R + pR;
      flux[im2] = rhoR * vx1R * vx2R;
      if (sd == 3) flux[im3] = rhoR * vx1R * vx3R;
      flux[ie] = (rhoeR + pR) * vx1R;
      for (int v = ia; v < nv; v++) flux[v] = rhoR * qR[v] * vx1R;
    } else {
      flux[0] = rhostar * ustar;
      flux[im1] = rhostar * ustar * ustar + pstar;
      flux[im2] = rhostar * ustar * (hp * vx2L + hm * vx2R);
      if (sd == 3) flux[im3] = rhostar * ustar * (hp * vx3L + hm * vx3R);
      flux[ie] = (hp * rhoestarL + hm * rhoestarR + pstar) * ustar;
      for (int v = ia; v < nv; v++) flux[v] = rhostar * ustar * (hp * qL[v] + hm * qR[v]);
    }
    return;
  }
  /* HLLC (Toro), source.F90:16124-16252 */
  double U[16], f[16], Us[16];
  if (ustar > 0.0) {
    U[0] = rhoL; U[im1] = rhoL * vx1L; U[im2] = rhoL * vx2L;
    if (sd == 3) U[im3] = rhoL * vx3L;
    U[ie] = rhoeL;
    for (int v = ia; v < nv; v++) U[v] = rhoL * qL[v];
    f[0] = rhoL * vx1L; f[im1] = rhoL * vx1L * vx1L + pL; f[im2] = rhoL * vx1L * vx2L;
    if (sd == 3) f[im3] = rhoL * vx1L * vx3L;
    f[ie] = (rhoeL + pL) * vx1L;
    for (int v = ia; v < nv; v++) f[v] = rhoL * vx1L * qL[v];
    if (sL >= 0.0) {
      for (int v = 0; v < nv; v++) flux[v] = f[v];
    } else {
      Us[0] = rhostarL; Us[im1] = rhostarL * ustar; Us[im2] = rhostarL * vx2L;
      if (sd == 3) Us[im3] = rhostarL * vx3L;
      Us[ie] = rhoestarL;
      for (int v = ia; v < nv; v++) Us[v] = rhostarL * qL[v];
      for (int v = 0; v < nv; v++) flux[v] = f[v] + sL * (Us[v] - U[v]);
    }
  } else {
    U[0] = rhoR; U[im1] = rhoR * vx1R; U[im2] = rhoR * vx2R;
    if (sd == 3) U[im3] = rhoR * vx3R;
    U[ie] = rhoeR;
    for (int v = ia; v < nv; v++) U[v] = rhoR * qR[v];
    f[0] = rhoR * vx1R; f[im1] = rhoR * vx1R * vx1R + pR; f[im2] = rhoR * vx1R * vx2R;
    if (sd == 3) f[im3] = rhoR * vx1R * vx3R;
    f[ie] = (rhoeR + pR) * vx1R;
    for (int v = ia; v < nv; v++) f[v] = rhoR * vx1R * qR[v];
    if (sR < 0.0) {
      for (int v = 0; v < nv; v++) flux[v] = f[v];
    } else {
      Us[0] = rhostarR; Us[im1] = rhostarR * ustar; Us[im2] = rhostarR * vx2R;
      if (sd == 3) Us[im3] = rhostarR * vx3R;
      Us[ie] = rhoestarR;
      for (int v = ia; v < nv; v++) Us[v] = rhostarR * qR[v];
      for (int v = 0; v < nv; v++) flux[v] = f[v] + sR * (Us[v] - U[v]);
    }
  }
}

/* HLLD / LHLLD (source.F90:13969-15024).  b = (b_normal, b_t1, [b_t2]). */
static void riemann_mhd_face(const rparams *P, const double *qL, const double *qR,
                             const double *bL, const double *bR, const double *gL,
                             const double *gR, double Tg, int hll, double *flux,
                             double *fluxb) {
  const int sd = P->sd, nv = P->nv;
  const int im1 = 1, im2 = 2, im3 = (sd == 3) ? 3 : 2, ie = im3 + 1, ia = 2 + sd;
  double rhoL = qL[0], inv_rhoL = 1.0 / rhoL;
  double vx1L = qL[1], vx2L = qL[2], vx3L = (sd == 3) ? qL[3] : 0.0;
  double v2L = vx1L * vx1L + vx2L * vx2L + vx3L * vx3L;
  double pL = qL[ie], eint, c2L, c2R;
  if (P->fasteos) {
    eint = pL / (gL[0] - 1.0);
    c2L = gL[1] * pL * inv_rhoL;
  } else if (P->helm) {
    double sound;
    face_helm(P, qL, Tg, &eint, &sound);
    c2L = sound * sound;
  } else {
    eint = pL * P->igmm1;
    c2L = P->gm * pL * inv_rhoL;
  }
  double rhoeL = eint + 0.5 * rhoL * v2L;
  double bx1L = bL[0], bx2L = bL[1], bx3L = (sd == 3) ? bL[2] : 0.0;
  double rhoR = qR[0], inv_rhoR = 1.0 / rhoR;
  double vx1R = qR[1], vx2R = qR[2], vx3R = (sd == 3) ? qR[3] : 0.0;
  double v2R = vx1R * vx1R + vx2R * vx2R + vx3R * vx3R;
  double pR = qR[ie];
  if (P->fasteos) {
    eint = pR / (gR[0] - 1.0);
    c2R = gR[1] * pR * inv_rhoR;
  } else if (P->helm) {
    double sound;
    face_helm(P, qR, Tg, &eint, &sound);
    c2R = sound * sound;
  } else {
    eint = pR * P->igmm1;
    c2R = P->gm * pR * inv_rhoR;
  }
  double rhoeR = eint + 0.5 * rhoR * v2R;
  double bx1R = bR[0], bx2R = bR[1], bx3R = (sd == 3) ? bR[2] : 0.0;

  double bx1 = 0.5 * (bx1L + bx1R), bx1_2 = bx1 * bx1;
  double pbL = 0.5 * (bx1_2 + bx2L * bx2L + bx3L * bx3L);
  double pbR = 0.5 * (bx1_2 + bx2R * bx2R + bx3R * bx3R);
  rhoeL = rhoeL + pbL;
  rhoeR = rhoeR + pbR;
  double ptL = pL + pbL, ptR = pR + pbR;
  double a2L = 2.0 * pbL * inv_rhoL, a2R = 2.0 * pbR * inv_rhoR;
  double an2L = bx1_2 * inv_rhoL, an2R = bx1_2 * inv_rhoR;
  double dummy = c2L + a2L;
  double cL = sqrt(0.5 * (c2L + a2L + sqrt(dummy * dummy - 4.0 * c2L * an2L)));
  dummy = c2R + a2R;
  double cR = sqrt(0.5 * (c2R + a2R + sqrt(dummy * dummy - 4.0 * c2R * an2R)));
  double max_c = fmax(cL, cR);
  double sL = fmin(vx1L, vx1R) - max_c;
  double sR = fmax(vx1L, vx1R) + max_c;
  double dsuL = sL - vx1L, dsuR = sR - vx1R;
  double inv_temp1 = 1.0 / (dsuR * rhoR - dsuL * rhoL);
  double ustar = (dsuR * rhoR * vx1R - dsuL * rhoL * vx1L - ptR + ptL) * inv_temp1;
  double dsumL = sL - ustar, dsumR = sR - ustar;
  double phi = 1.0;
  if (P->flux == ORC_LHLLD) {
    /* source.F90:14364-14384 */
    dummy = v2L + a2L;
    double cu2L = 0.5 * (v2L + a2L + sqrt(dummy * dummy - 4.0 * v2L * an2L));
    dummy = v2R + a2R;
    double cu2R = 0.5 * (v2R + a2R + sqrt(dummy * dummy - 4.0 * v2R * an2R));
    double machL = sqrt(cu2L) / cL, machR = sqrt(cu2R) / cR;
    double mx = fmax(machL, machR);
    if (mx < 0.6) {
      double chi = fmin(1.0, fmax(machL, machR));
      chi = fmax(P->mach_cutoff, chi);
      phi = chi * (2.0 - chi);
    } else {
      phi = 1.0;
    }
  }
  if (hll) {
    /* USE_SHOCK_FLATTENING: two-wave HLL flux on a flattened face
     * (source.F90:14386-14486) */
    double UL[16], UR[16], fL[16], fR[16], UbL[3], UbR[3], fbL[3], fbR[3];
    sL = fmin(0.0, sL);
    sR = fmax(0.0, sR);
    UL[0] = rhoL; UL[im1] = rhoL * vx1L; UL[im2] = rhoL * vx2L;
    if (sd == 3) UL[im3] = rhoL * vx3L;
    UL[ie] = rhoeL;
    for (int v = ia; v < nv; v++) UL[v] = rhoL * qL[v];
    UbL[0] = bx1; UbL[1] = bx2L; UbL[2] = bx3L;
    UR[0] = rhoR; UR[im1] = rhoR * vx1R; UR[im2] = rhoR * vx2R;
    if (sd == 3) UR[im3] = rhoR * vx3R;
    UR[ie] = rhoeR;
    for (int v = ia; v < nv; v++) UR[v] = rhoR * qR[v];
    UbR[0] = bx1; UbR[1] = bx2R; UbR[2] = bx3R;
    fL[0] = rhoL * vx1L;
    fL[im1] = rhoL * vx1L * vx1L + ptL - bx1_2;
    fL[im2] = rhoL * vx1L * vx2L - bx1 * bx2L;
    if (sd == 3) fL[im3] = rhoL * vx1L * vx3L - bx1 * bx3L;
    fL[ie] = (rhoeL + ptL) * vx1L - bx1 * (bx1 * vx1L + bx2L * vx2L + bx3L * vx3L);
    for (int v = ia; v < nv; v++) fL[v] = UL[v] * vx1L;
    fbL[0] = 0.0; fbL[1] = vx1L * bx2L - vx2L * bx1; fbL[2] = vx1L * bx3L - vx3L * bx1;
    fR[0] = rhoR * vx1R;
    fR[im1] = rhoR * vx1R * vx1R + ptR - bx1_2;
    fR[im2] = rhoR * vx1R * vx2R - bx1 * bx2R;
    if (sd == 3) fR[im3] = rhoR * vx1R * vx3R - bx1 * bx3R;
    fR[ie] = (rhoeR + ptR) * vx1R - bx1 * (bx1 * vx1R + bx2R * vx2R + bx3R * vx3R);
    for (int v = ia; v < nv; v++) fR[v] = UR[v] * vx1R;
    fbR[0] = 0.0; fbR[1] = vx1R * bx2R - vx2R * bx1; fbR[2] = vx1R * bx3R - vx3R * bx1;
    double dm = 1.0 / (sR - sL), ph = sL * sR;
    for (int v = 0; v < nv; v++) flux[v] = dm * (sR * fL[v] - sL * fR[v] + ph * (UR[v] - UL[v]));
    for (int v = 0; v < sd; v++)
      fluxb[v] = dm * (sR * fbL[v] - sL * fbR[v] + ph * (UbR[v] - UbL[v]));
    return;
  }
  double pstar = (dsuR * rhoR * ptL - dsuL * rhoL * ptR - phi * rhoL * rhoR * dsuR * dsuL * (vx1L - vx1R)) * inv_temp1;
  double rhostarL = rhoL * dsuL / dsumL;
  double rhostarR = rhoR * dsuR / dsumR;
  double a1_inv = 1.0 / (rhoL * dsuL * dsumL - bx1_2);
  double vx2starL = vx2L - bx1 * bx2L * (ustar - vx1L) * a1_inv;
  double bx2starL = bx2L * (rhoL * dsuL * dsuL - bx1_2) * a1_inv;
  double vx3starL = vx3L - bx1 * bx3L * (ustar - vx1L) * a1_inv;
  double bx3starL = bx3L * (rhoL * dsuL * dsuL - bx1_2) * a1_inv;
  double a2_inv = 1.0 / (rhoR * dsuR * dsumR - bx1_2);
  double vx2starR = vx2R - bx1 * bx2R * (ustar - vx1R) * a2_inv;
  double bx2starR = bx2R * (rhoR * dsuR * dsuR - bx1_2) * a2_inv;
  double vx3starR = vx3R - bx1 * bx3R * (ustar - vx1R) * a2_inv;
  double bx3starR = bx3R * (rhoR * dsuR * dsuR - bx1_2) * a2_inv;
  double vdotbL = vx1L * bx1 + vx2L * bx2L + vx3L * bx3L;
  double vdotbR = vx1R * bx1 + vx2R * bx2R + vx3R * bx3R;
  double vdotbstarL = ustar * bx1 + vx2starL * bx2starL + vx3starL * bx3starL;
  double vdotbstarR = ustar * bx1 + vx2starR * bx2starR + vx3starR * bx3starR;
  double rhoestarL = (dsuL * rhoeL - ptL * vx1L + pstar * ustar + bx1 * (vdotbL - vdotbstarL)) / dsumL;
  double rhoestarR = (dsuR * rhoeR - ptR * vx1R + pstar * ustar + bx1 * (vdotbR - vdotbstarR)) / dsumR;
  double sqL = sqrt(rhostarL), sqR = sqrt(rhostarR);
  double abs_bx1 = fabs(bx1);
  double sstarL = ustar - abs_bx1 / sqL;
  double sstarR = ustar + abs_bx1 / sqR;
  double inv_sqr_bar = 1.0 / (sqL + sqR);
  double sgn = copysign(1.0, bx1);
  double vx22star = (sqL * vx2starL + sqR * vx2starR + (bx2starR - bx2starL) * sgn) * inv_sqr_bar;
  double bx22star = (sqL * bx2starR + sqR * bx2starL + sqL * sqR * (vx2starR - vx2starL) * sgn) * inv_sqr_bar;
  double vx32star = (sqL * vx3starL + sqR * vx3starR + (bx3starR - bx3starL) * sgn) * inv_sqr_bar;
  double bx32star = (sqL * bx3starR + sqR * bx3starL + sqL * sqR * (vx3starR - vx3starL) * sgn) * inv_sqr_bar;
  double vdotb2star = ustar * bx1 + vx22star * bx22star + vx32star * bx32star;
  double rhoe2starL = rhoestarL - sqL * (vdotbstarL - vdotb2star) * sgn;
  double rhoe2starR = rhoestarR + sqR * (vdotbstarR - vdotb2star) * sgn;

  if (P->flux == ORC_LHLLD) {
    double rho, vx1, vx2, vx3, rhoe, bx2, bx3, pres;
    if (P->all_mach && sL > 0.0) {
      rho = rhoL; vx1 = vx1L; vx2 = vx2L; vx3 = vx3L; rhoe = rhoeL; bx2 = bx2L; bx3 = bx3L; pres = ptL;
    } else if (sstarL > 0.0) {
      rho = rhostarL; vx1 = ustar; vx2 = vx2starL; vx3 = vx3starL; rhoe = rhoestarL; bx2 = bx2starL; bx3 = bx3starL; pres = pstar;
    } else if (ustar > 0.0) {
      rho = rhostarL; vx1 = ustar; vx2 = vx22star; vx3 = vx32star; rhoe = rhoe2starL; bx2 = bx22star; bx3 = bx32star; pres = pstar;
    } else if (sstarR > 0.0) {
      rho = rhostarR; vx1 = ustar; vx2 = vx22star; vx3 = vx32star; rhoe = rhoe2starR; bx2 = bx22star; bx3 = bx32star; pres = pstar;
    } else if (!P->all_mach || sR > 0.0) {
      rho = rhostarR; vx1 = ustar; vx2 = vx2starR; vx3 = vx3starR; rhoe = rhoestarR; bx2 = bx2starR; bx3 = bx3starR; pres = pstar;
    } else {
      rho = rhoR; vx1 = vx1R; vx2 = vx2R; vx3 = vx3R; rhoe = rhoeR; bx2 = bx2R; bx3 = bx3R; pres = ptR;
    }
    flux[0] = rho * vx1;
    flux[im1] = rho * vx1 * vx1 + pres - bx1_2;
    flux[im2] = rho * vx1 * vx2 - bx1 * bx2;
    if (sd == 3) flux[im3] = rho * vx1 * vx3 - bx1 * bx3;
    flux[ie] = (rhoe + pres) * vx1 - bx1 * (bx1 * vx1 + vx2 * bx2 + vx3 * bx3);
    if (nv > ia) {
      double s2 = copysign(1.0, vx1), hp = 0.5 * (1.0 + s2), hm = 0.5 * (1.0 - s2);
      for (int v = ia; v < nv; v++) flux[v] = rho * vx1 * (hp * qL[v] + hm * qR[v]);
    }
    fluxb[0] = 0.0;
    fluxb[1] = vx1 * bx2 - bx1 * vx2;
    if (sd == 3) fluxb[2] = vx1 * bx3 - bx1 * vx3;
    return;
  }
  /* HLLD (source.F90:14736-15016) */
  double U[16], f[16], Us[16], Uss[16], Ub[3], fb[3], Ubs[3], Ubss[3];
  if (ustar > 0.0) {
    U[0] = rhoL; U[im1] = rhoL * vx1L; U[im2] = rhoL * vx2L;
    if (sd == 3) U[im3] = rhoL * vx3L;
    U[ie] = rhoeL;
    for (int v = ia; v < nv; v++) U[v] = rhoL * qL[v];
    Ub[0] = bx1; Ub[1] = bx2L; Ub[2] = bx3L;
    f[0] = rhoL * vx1L;
    f[im1] = rhoL * vx1L * vx1L + ptL - bx1_2;
    f[im2] = rhoL * vx1L * vx2L - bx1 * bx2L;
    if (sd == 3) f[im3] = rhoL * vx1L * vx3L - bx1 * bx3L;
    f[ie] = (rhoeL + ptL) * vx1L - bx1 * vdotbL;
    for (int v = ia; v < nv; v++) f[v] = rhoL * vx1L * qL[v];
    fb[0] = 0.0;
    fb[1] = vx1L * bx2L - bx1 * vx2L;
    fb[2] = vx1L * bx3L - bx1 * vx3L;
    if (sL > 0.0) {
      for (int v = 0; v < nv; v++) flux[v] = f[v];
      for (int v = 0; v < sd; v++) fluxb[v] = fb[v];
    } else {
      Us[0] = rhostarL; Us[im1] = rhostarL * ustar; Us[im2] = rhostarL * vx2starL;
      if (sd == 3) Us[im3] = rhostarL * vx3starL;
      Us[ie] = rhoestarL;
      for (int v = ia; v < nv; v++) Us[v] = rhostarL * qL[v];
      Ubs[0] = bx1; Ubs[1] = bx2starL; Ubs[2] = bx3starL;
      if (sstarL > 0.0) {
        for (int v = 0; v < nv; v++) flux[v] = f[v] + sL * (Us[v] - U[v]);
        for (int v = 0; v < sd; v++) fluxb[v] = fb[v] + sL * (Ubs[v] - Ub[v]);
      } else {
        Uss[0] = rhostarL; Uss[im1] = rhostarL * ustar; Uss[im2] = rhostarL * vx22star;
        if (sd == 3) Uss[im3] = rhostarL * vx32star;
        Uss[ie] = rhoe2starL;
        for (int v = ia; v < nv; v++) Uss[v] = rhostarL * qL[v];
        Ubss[0] = bx1; Ubss[1] = bx22star; Ubss[2] = bx32star;
        for (int v = 0; v < nv; v++) flux[v] = f[v] + sL * (Us[v] - U[v]) + sstarL * (Uss[v] - Us[v]);
        for (int v = 0; v < sd; v++) fluxb[v] = fb[v] + sL * (Ubs[v] - Ub[v]) + sstarL * (Ubss[v] - Ubs[v]);
      }
    }
  } else {
    U[0] = rhoR; U[im1] = rhoR * vx1R; U[im2] = rhoR * vx2R;
    if (sd == 3) U[im3] = rhoR * vx3R;
    U[ie] = rhoeR;
    for (int v = ia; v < nv; v++) U[v] = rhoR * qR[v];
    Ub[0] = bx1; Ub[1] = bx2R; Ub[2] = bx3R;
    f[0] = rhoR * vx1R;
    f[im1] = rhoR * vx1R * vx1R + ptR - bx1_2;
    f[im2] = rhoR * vx1R * vx2R - bx1 * bx2R;
    if (sd == 3) f[im3] = rhoR * vx1R * vx3R - bx1 * bx3R;
    f[ie] = (rhoeR + ptR) * vx1R - bx1 * vdotbR;
    for (int v = ia; v < nv; v++) f[v] = rhoR * vx1R * qR[v];
    fb[0] = 0.0;
    fb[1] = vx1R * bx2R - bx1 * vx2R;
    fb[2] = vx1R * bx3R - bx1 * vx3R;
    if (sR < 0.0) {
      for (int v = 0; v < nv; v++) flux[v] = f[v];
      for (int v = 0; v < sd; v++) fluxb[v] = fb[v];
    } else {
      Us[0] = rhostarR; Us[im1] = rhostarR * ustar; Us[im2] = rhostarR * vx2starR;
      if (sd == 3) Us[im3] = rhostarR * vx3starR;
      Us[ie] = rhoestarR;
      for (int v = ia; v < nv; v++) Us[v] = rhostarR * qR[v];
      Ubs[0] = bx1; Ubs[1] = bx2starR; Ubs[2] = bx3starR;
      if (sstarR < 0.0) {
        for (int v = 0; v < nv; v++) flux[v] = f[v] + sR * (Us[v] - U[v]);
        for (int v = 0; v < sd; v++) fluxb[v] = fb[v] + sR * (Ubs[v] - Ub[v]);
      } else {
        Uss[0] = rhostarR; Uss[im1] = rhostarR * ustar; Uss[im2] = rhostarR * vx22star;
        if (sd == 3) Uss[im3] = rhostarR * vx32star;
        Uss[ie] = rhoe2starR;
        for (int v = ia; v < nv; v++) Uss[v] = rhostarR * qR[v];
        Ubss[0] = bx1; Ubss[1] = bx22star; Ubss[2] = bx32star;
        for (int v = 0; v < nv; v++) flux[v] = f[v] + sR * (Us[v] - U[v]) + sstarR * (Uss[v] - Us[v]);
        for (int v = 0; v < sd; v++) fluxb[v] = fb[v] + sR * (Ubs[v] - Ub[v]) + sstarR * (Ubss[v] - Ubs[v]);
      }
    }
  }
}

void orc_riemann_hydro(const orc_config *cfg, int n, const double *qL,
                       const double *qR, double *flux) {
  rparams P = make_rp(cfg);
  for (int i = 0; i < n; i++)
    riemann_hydro_face(&P, qL + (long)i * P.nv, qR + (long)i * P.nv, NULL, NULL, 0.0, 0,
                       flux + (long)i * P.nv);
}

void orc_riemann_mhd(const orc_config *cfg, int n, const double *qL,
                     const double *qR, const double *bL, const double *bR,
                     double *flux, double *fluxb) {
  rparams P = make_rp(cfg);
  for (int i = 0; i < n; i++)
    riemann_mhd_face(&P, qL + (long)i * P.nv, qR + (long)i * P.nv, bL + (long)i * P.sd,
                     bR + (long)i * P.sd, NULL, NULL, 0.0, 0, flux + (long)i * P.nv,
                     fluxb + (long)i * P.sd);
}

/* -------------------------------------------------------------- sweeps */

/* Equilibrium pressure of a (possibly ghost) x2 row: the row the boundary
 * map sends it to (periodic wrap or mirror). */
static double eq_p_row(const orc_grid *g, int j) {
  const int n = g->n[1];
  if (j < 0) j = (g->c.bc[2] == ORC_BC_PERIODIC) ? j + n : -1 - j;
  else if (j >= n) j = (g->c.bc[3] == ORC_BC_PERIODIC) ? j - n : 2 * n - 1 - j;
  return g->eq_p_cc[j];
}

/* USE_SHOCK_FLATTENING detector (source.F90:8073-8131): a cell whose summed
 * relative pressure jump sum_d |p(c+e_d) - p(c-e_d)| / p(c) exceeds eps_sf
 * flags itself and its 3^sdims neighbourhood; flags cover cells -1..n.  With
 * USE_WB the full pressure is prim + eq; the equilibrium of a ghost row is
 * that of the row its boundary map copies (parity unpinned there: the
 * reference's ghost-row equilibrium comes from the application). */
static void shock_flags(orc_grid *g) {
  const int sd = g->sd, nv = g->nv, ip = g->ip;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const int k0 = (sd == 3) ? -1 : 0, k1 = (sd == 3) ? n3 : 0;
  for (int k = k0; k <= k1; k++)
    for (int j = -1; j <= n2; j++)
      for (int i = -1; i <= n1; i++) g->flat[cidx(g, i, j, k)] = 0;
  const int kd0 = (sd == 3) ? -2 : 0, kd1 = (sd == 3) ? n3 + 1 : 0;
  const int wb = g->c.use_wb;
  for (int k = kd0; k <= kd1; k++)
    for (int j = -2; j <= n2 + 1; j++)
      for (int i = -2; i <= n1 + 1; i++) {
        double pc = g->prim[cidx(g, i, j, k) * nv + ip];
        double pjp = g->prim[cidx(g, i, j + 1, k) * nv + ip];
        double pjm = g->prim[cidx(g, i, j - 1, k) * nv + ip];
        double pip = g->prim[cidx(g, i + 1, j, k) * nv + ip];
        double pim = g->prim[cidx(g, i - 1, j, k) * nv + ip];
        double pkp = 0.0, pkm = 0.0;
        if (sd == 3) {
          pkp = g->prim[cidx(g, i, j, k + 1) * nv + ip];
          pkm = g->prim[cidx(g, i, j, k - 1) * nv + ip];
        }
        if (wb) {
          const double e = eq_p_row(g, j);
          pc = pc + e;
          pjp = pjp + eq_p_row(g, j + 1);
          pjm = pjm + eq_p_row(g, j - 1);
          pip = pip + e;
          pim = pim + e;
          if (sd == 3) {
            pkp = pkp + e;
            pkm = pkm + e;
          }
        }
        double tmp;
        if (sd == 3)
          tmp = (fabs(pkp - pkm) + fabs(pjp - pjm) + fabs(pip - pim)) / pc;
        else
          tmp = (fabs(pjp - pjm) + fabs(pip - pim)) / pc;
        if (tmp > g->c.eps_sf) {
          for (int dk = (sd == 3 ? -1 : 0); dk <= (sd == 3 ? 1 : 0); dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) g->flat[cidx(g, i + di, j + dj, k + dk)] = 1;
        }
      }
}

/* The shock-flattening flags of the current primitive state, as a stage
 * computes them (WB deviation, halos, boundaries, detector), written to out
 * as the flag box of cells -1..n on each axis (x3 only in 3D), x1 fastest.
 * The primitive array is restored afterwards.  Test accessor. */
int orc_shock_flags(orc_grid *g, unsigned char *out) {
  if (!g->c.shock_flattening) return -1;
  const int sd = g->sd, nv = g->nv, ip = g->ip;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const size_t bytes = sizeof(double) * (size_t)g->e1 * g->e2 * g->e3 * nv;
  double *save = (double *)malloc(bytes);
  if (!save) return -2;
  memcpy(save, g->prim, bytes);
  if (g->c.use_wb)
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          double *p = g->prim + cidx(g, i, j, k) * nv;
          p[0] = p[0] - g->eq_rho_cc[j];
          p[ip] = p[ip] - g->eq_p_cc[j];
        }
  halo_periodic(g, g->prim, nv);
  physical_bc(g);
  shock_flags(g);
  const int k0 = (sd == 3) ? -1 : 0, k1 = (sd == 3) ? n3 : 0;
  long o = 0;
  for (int k = k0; k <= k1; k++)
    for (int j = -1; j <= n2; j++)
      for (int i = -1; i <= n1; i++) out[o++] = g->flat[cidx(g, i, j, k)];
  memcpy(g->prim, save, bytes);
  free(save);
  return 0;
}

/* One directional sweep over all interior pencils (x1 :7960-9105, x2
 * :9107-10301, x3 :10303-11364): reconstruct, add the face equilibrium
 * (WB), rotate into the Riemann frame, solve, rotate back, record EMFs and
 * contact signs, subtract p_eq from the normal momentum flux, add the ETOT
 * potential flux and difference into res (x1 overwrites, x2/x3 add). */
static void sweep(orc_grid *g, int d) {
  const int sd = g->sd, nv = g->nv, ng = g->ng, mhd = g->mhd;
  const int n = g->n[d];
  const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
  const int fasteos = g->c.fasteos;
  const double inv_dl = g->inv_dx[d];
  rparams P = make_rp(&g->c);
  const int helm_face = TABULATED(g) && !fasteos;
  if (helm_face) P.helm = g->helm;
  long stride = (d == 0) ? 1 : (d == 1) ? g->e1 : (long)g->e1 * g->e2;
  const int ip = g->ip, ias1 = g->ias1;
  const int imd = g->iv[d];

#pragma omp parallel
  {
    double *qL = (double *)malloc(sizeof(double) * (n + 1) * nv);
    double *qR = (double *)malloc(sizeof(double) * (n + 1) * nv);
    double *fx = (double *)malloc(sizeof(double) * (n + 1) * nv);
    double *bL = (double *)malloc(sizeof(double) * (n + 1) * 3);
    double *bR = (double *)malloc(sizeof(double) * (n + 1) * 3);
    double *fb = (double *)malloc(sizeof(double) * (n + 1) * 3);
    double *gL = (double *)malloc(sizeof(double) * (n + 1) * 2);
    double *gR = (double *)malloc(sizeof(double) * (n + 1) * 2);
    int np = g->n[t1] * g->n[t2];
#pragma omp for schedule(static)
    for (int pidx = 0; pidx < np; pidx++) {
      int a = pidx % g->n[t1], b = pidx / g->n[t1];
      int id[3];
      id[t1] = a;
      id[t2] = b;
      id[d] = 0;
      long c0 = cidx(g, id[0], id[1], id[2]);
      const double *pr = g->prim, *bc = g->bcc, *gf = g->gamf;
      /* reconstruction of cells -1..n (faces 0..n) */
      for (int c = -1; c <= n; c++) {
        long cm = c0 + (long)(c - 1) * stride, cc = c0 + (long)c * stride, cp = c0 + (long)(c + 1) * stride;
        /* LIM5TH: two more cells (ng = 3); the indices stay inside the ghosts */
        /* a flattened cell takes the LIM2ND slope whatever the scheme
         * (x1 :8743-8869, x2 :9635-9761, x3 :10765-10891) */
        const int fl = g->c.shock_flattening && g->flat[cc];
        const int r5 = g->c.recon == ORC_LIM5TH && !fl;
        const int l2 = g->c.recon == ORC_LIM2ND || fl;
        long cmm = r5 ? c0 + (long)(c - 2) * stride : cm, cpp = r5 ? c0 + (long)(c + 2) * stride : cp;
        double l, r;
        for (int v = 0; v < nv; v++) {
          double qm = pr[cm * nv + v], qc = pr[cc * nv + v], qp = pr[cp * nv + v];
          if (r5)
            rec_lim5(pr[cmm * nv + v], qm, qc, qp, pr[cpp * nv + v], &l, &r);
          else if (l2)
            rec_lim2nd(qm, qc, qp, &l, &r);
          else if (v <= ip)
            rec_pph(qm, qc, qp, &l, &r);
          else
            rec_pph_lim(qm, qc, qp, &l, &r);
          if (c >= 0) qR[(long)c * nv + v] = l;
          if (c <= n - 1) qL[(long)(c + 1) * nv + v] = r;
        }
        if (mhd) {
          for (int comp = 0; comp < sd; comp++) {
            if (comp == d) continue;
            double qm = bc[cm * 3 + comp], qc = bc[cc * 3 + comp], qp = bc[cp * 3 + comp];
            if (r5)
              rec_lim5(bc[cmm * 3 + comp], qm, qc, qp, bc[cpp * 3 + comp], &l, &r);
            else if (l2)
              rec_lim2nd(qm, qc, qp, &l, &r);
            else
              rec_pph(qm, qc, qp, &l, &r);
            if (c >= 0) bR[(long)c * 3 + comp] = l;
            if (c <= n - 1) bL[(long)(c + 1) * 3 + comp] = r;
          }
        }
        if (fasteos) {
          for (int comp = 0; comp < 2; comp++) {
            double qm = gf[cm * 2 + comp], qc = gf[cc * 2 + comp], qp = gf[cp * 2 + comp];
            if (r5)
              rec_lim5(gf[cmm * 2 + comp], qm, qc, qp, gf[cpp * 2 + comp], &l, &r);
            else if (l2)
              rec_lim2nd(qm, qc, qp, &l, &r);
            else
              rec_pph(qm, qc, qp, &l, &r);
            if (c >= 0) gR[(long)c * 2 + comp] = l;
            if (c <= n - 1) gL[(long)(c + 1) * 2 + comp] = r;
          }
        }
      }
      /* normal B from the face field */
      if (mhd)
        for (int f = 0; f <= n; f++) {
          int fi[3];
          fi[t1] = a; fi[t2] = b; fi[d] = f;
          double bn = g->bf[d][fidx(g, d, fi[0], fi[1], fi[2])];
          bL[(long)f * 3 + d] = bn;
          bR[(long)f * 3 + d] = bn;
        }
      /* rotation into the Riemann frame (x2 :9716-9727, x3 :10845-10864) */
      for (int f = 0; f <= n; f++) {
        for (int s = 0; s < 2; s++) {
          double *q = (s == 0 ? qL : qR) + (long)f * nv;
          double *bb = (s == 0 ? bL : bR) + (long)f * 3;
          if (d == 1) {
            double v1 = q[1], v2 = q[2];
            q[1] = v2;
            q[2] = -v1;
            if (mhd) {
              double b1 = bb[0], b2 = bb[1];
              bb[0] = b2;
              bb[1] = -b1;
            }
          } else if (d == 2) {
            double v1 = q[1], v2 = q[2], v3 = q[3];
            q[1] = v3;
            q[2] = v1;
            q[3] = v2;
            if (mhd) {
              double b1 = bb[0], b2 = bb[1], b3 = bb[2];
              bb[0] = b3;
              bb[1] = b1;
              bb[2] = b2;
            }
          }
        }
      }
      /* WB: add face equilibrium (x1 :8473-8486) */
      if (g->c.use_wb)
        for (int f = 0; f <= n; f++) {
          int jy = (d == 1) ? f : (d == 0 ? a : b); /* x2 index of this face */
          double rho_eq = (d == 1) ? g->eq_rho_f[jy] : g->eq_rho_cc[jy];
          double p_eq = (d == 1) ? g->eq_p_f[jy] : g->eq_p_cc[jy];
          qL[(long)f * nv] = qL[(long)f * nv] + rho_eq;
          qR[(long)f * nv] = qR[(long)f * nv] + rho_eq;
          qL[(long)f * nv + ip] = qL[(long)f * nv + ip] + p_eq;
          qR[(long)f * nv + ip] = qR[(long)f * nv + ip] + p_eq;
          if (fasteos) {
            /* x1 :8715-8727, x2 :10038-10048, x3 :11180-11190 */
            double ge_eq = (d == 1) ? g->eq_ge_f[jy] : g->eq_ge_cc[jy];
            gL[(long)f * 2] = gL[(long)f * 2] + ge_eq;
            gR[(long)f * 2] = gR[(long)f * 2] + ge_eq;
          }
        }
      /* Riemann; without FASTEOS the face T guess is the mean of the two
       * cells' T (x1 :8488-8494, x2 :9766-9772, x3 :10908-10914) */
      for (int f = 0; f <= n; f++) {
        double Tg = 0.0;
        if (helm_face)
          Tg = 0.5 * (g->temp[c0 + (long)(f - 1) * stride] + g->temp[c0 + (long)f * stride]);
        /* a face beside a flattened cell takes the HLL flux (x1 :8731-8740) */
        const int hll = g->c.shock_flattening &&
                        (g->flat[c0 + (long)(f - 1) * stride] || g->flat[c0 + (long)f * stride]);
        if (mhd)
          riemann_mhd_face(&P, qL + (long)f * nv, qR + (long)f * nv, bL + (long)f * 3,
                           bR + (long)f * 3, gL + (long)f * 2, gR + (long)f * 2, Tg, hll,
                           fx + (long)f * nv, fb + (long)f * 3);
        else
          riemann_hydro_face(&P, qL + (long)f * nv, qR + (long)f * nv, gL + (long)f * 2,
                             gR + (long)f * 2, Tg, hll, fx + (long)f * nv);
      }
      for (int f = 0; f <= n; f++) {
        double *F = fx + (long)f * nv, *FB = fb + (long)f * 3;
        /* back-rotation (x2 :10101-10111, x3 :11241-11253) */
        if (d == 1) {
          double v1 = F[1], v2 = F[2];
          F[1] = -v2;
          F[2] = v1;
          if (mhd) {
            double b1 = FB[0], b2 = FB[1];
            FB[0] = -b2;
            FB[1] = b1;
          }
        } else if (d == 2) {
          double v1 = F[1], v2 = F[2], v3 = F[3];
          F[1] = v2;
          F[2] = v3;
          F[3] = v1;
          if (mhd) {
            double b1 = FB[0], b2 = FB[1], b3 = FB[2];
            FB[0] = b2;
            FB[1] = b3;
            FB[2] = b1;
          }
        }
        if (mhd) {
          /* face EMFs and contact-upwind sign (x1 :8926-8951) */
          int fi[3];
          fi[t1] = a; fi[t2] = b; fi[d] = f;
          long e = gidx(g, d, fi[0], fi[1], fi[2]);
          double *E = g->emf[d] + e * 3;
          if (d == 0) {
            if (sd == 3) { E[0] = 0.0; E[1] = FB[2]; }
            E[2] = -FB[1];
          } else if (d == 1) {
            if (sd == 3) { E[0] = -FB[2]; E[1] = 0.0; }
            E[2] = FB[0];
          } else {
            E[0] = FB[1];
            E[1] = -FB[0];
            E[2] = 0.0;
          }
          double tmp = 0.5 * (qL[(long)f * nv] + qR[(long)f * nv]) * g->smax;
          double mach = F[0] / tmp;
          double s = 0.0;
          if (mach > g->c.mach_ct) s = 1.0;
          if (mach < -g->c.mach_ct) s = -1.0;
          g->sgn[d][e] = s;
        }
        if (g->c.use_wb) {
          int jy = (d == 1) ? f : (d == 0 ? a : b);
          double p_eq = (d == 1) ? g->eq_p_f[jy] : g->eq_p_cc[jy];
          F[imd] = F[imd] - p_eq;
        }
        if (g->c.evolve_etot) {
          int jy = (d == 1) ? f : (d == 0 ? a : b);
          double phi = (d == 1) ? g->phi_f[jy] : g->phi_cc[jy];
          F[ip] = F[ip] + F[0] * phi;
        }
      }
      /* flux divergence (x1 :9036-9102 overwrites; x2/x3 accumulate) */
      for (int c = 0; c < n; c++) {
        int ci[3];
        ci[t1] = a; ci[t2] = b; ci[d] = c;
        double *R = g->res + iidx(g, ci[0], ci[1], ci[2]) * nv;
        const double *Fm = fx + (long)c * nv, *Fp = fx + (long)(c + 1) * nv;
        if (d == 0)
          for (int v = 0; v < nv; v++) R[v] = (Fp[v] - Fm[v]) * inv_dl;
        else
          for (int v = 0; v < nv; v++) R[v] = R[v] + (Fp[v] - Fm[v]) * inv_dl;
      }
    }
    free(qL); free(qR); free(fx); free(bL); free(bR); free(fb); free(gL); free(gR);
  }
  (void)ng;
  (void)ias1;
}

/* ----------------------------------------------- constrained transport */

static inline double cell_emf(const orc_grid *g, int comp, int i, int j, int k) {
  /* -(v x B)_comp from cell-centred v and b_cc (source.F90:11974-11985) */
  long c = cidx(g, i, j, k);
  const double *p = g->prim + c * g->nv, *b = g->bcc + c * 3;
  if (comp == 2) return -p[g->iv[0]] * b[1] + p[g->iv[1]] * b[0];
  if (comp == 0) return p[g->iv[2]] * b[1] - p[g->iv[1]] * b[2];
  return p[g->iv[0]] * b[2] - p[g->iv[2]] * b[0];
}

/* Gardiner-Stone CT-contact corner EMF (source.F90:11966-12027) */
static inline double corner(double eS, double eW, double eN, double eE, double eSW,
                            double eNW, double eNE, double eSE, double sS, double sN,
                            double sW, double sE) {
  double dS = 0.5 * (1.0 + sS) * (eW - eSW) + 0.5 * (1.0 - sS) * (eE - eSE);
  double dN = 0.5 * (1.0 + sN) * (eNW - eW) + 0.5 * (1.0 - sN) * (eNE - eE);
  double dW = 0.5 * (1.0 + sW) * (eS - eSW) + 0.5 * (1.0 - sW) * (eN - eNW);
  double dE = 0.5 * (1.0 + sE) * (eSE - eS) + 0.5 * (1.0 - sE) * (eNE - eN);
  return 0.25 * (eS + eW + eN + eE + dS - dN + dW - dE);
}

static void ct_update(orc_grid *g, double a1, double a2, double a3) {
  const int sd = g->sd;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  double *E1 = g->emf[0], *E2 = g->emf[1], *E3 = g->emf[2];
  double *S1 = g->sgn[0], *S2 = g->sgn[1], *S3 = g->sgn[2];
  /* emfx3_cor */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; k++)
    for (int j = 0; j <= n2; j++)
      for (int i = 0; i <= n1; i++) {
        double eN = E1[gidx(g, 0, i, j, k) * 3 + 2];
        double eE = E2[gidx(g, 1, i, j, k) * 3 + 2];
        double eS = E1[gidx(g, 0, i, j - 1, k) * 3 + 2];
        double eW = E2[gidx(g, 1, i - 1, j, k) * 3 + 2];
        double eSW = cell_emf(g, 2, i - 1, j - 1, k), eNW = cell_emf(g, 2, i - 1, j, k);
        double eNE = cell_emf(g, 2, i, j, k), eSE = cell_emf(g, 2, i, j - 1, k);
        g->cor[2][kidx(g, 2, i, j, k)] =
            corner(eS, eW, eN, eE, eSW, eNW, eNE, eSE, S1[gidx(g, 0, i, j - 1, k)],
                   S1[gidx(g, 0, i, j, k)], S2[gidx(g, 1, i - 1, j, k)], S2[gidx(g, 1, i, j, k)]);
      }
  if (sd == 3) {
    /* emfx1_cor (source.F90:12050-12107) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= n3; k++)
      for (int j = 0; j <= n2; j++)
        for (int i = 0; i < n1; i++) {
          double eN = E2[gidx(g, 1, i, j, k) * 3 + 0];
          double eE = E3[gidx(g, 2, i, j, k) * 3 + 0];
          double eS = E2[gidx(g, 1, i, j, k - 1) * 3 + 0];
          double eW = E3[gidx(g, 2, i, j - 1, k) * 3 + 0];
          double eSW = cell_emf(g, 0, i, j - 1, k - 1), eNW = cell_emf(g, 0, i, j - 1, k);
          double eNE = cell_emf(g, 0, i, j, k), eSE = cell_emf(g, 0, i, j, k - 1);
          g->cor[0][kidx(g, 0, i, j, k)] =
              corner(eS, eW, eN, eE, eSW, eNW, eNE, eSE, S2[gidx(g, 1, i, j, k - 1)],
                     S2[gidx(g, 1, i, j, k)], S3[gidx(g, 2, i, j - 1, k)], S3[gidx(g, 2, i, j, k)]);
        }
    /* emfx2_cor (source.F90:12128-12185) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i <= n1; i++) {
          double eN = E3[gidx(g, 2, i, j, k) * 3 + 1];
          double eE = E1[gidx(g, 0, i, j, k) * 3 + 1];
          double eS = E3[gidx(g, 2, i - 1, j, k) * 3 + 1];
          double eW = E1[gidx(g, 0, i, j, k - 1) * 3 + 1];
          double eSW = cell_emf(g, 1, i - 1, j, k - 1), eNW = cell_emf(g, 1, i, j, k - 1);
          double eNE = cell_emf(g, 1, i, j, k), eSE = cell_emf(g, 1, i - 1, j, k);
          g->cor[1][kidx(g, 1, i, j, k)] =
              corner(eS, eW, eN, eE, eSW, eNW, eNE, eSE, S3[gidx(g, 2, i - 1, j, k)],
                     S3[gidx(g, 2, i, j, k)], S1[gidx(g, 0, i, j, k - 1)], S1[gidx(g, 0, i, j, k)]);
        }
  }
  const double idx1 = g->inv_dx[0], idx2 = g->inv_dx[1], idx3 = g->inv_dx[2];
  double *C1 = g->cor[0], *C2 = g->cor[1], *C3 = g->cor[2];
  /* face-B update (source.F90:12208-12435) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i <= n1; i++) {
        double tmp;
        if (sd == 3)
          tmp = -(C2[kidx(g, 1, i, j, k + 1)] - C2[kidx(g, 1, i, j, k)]) * idx3 +
                (C3[kidx(g, 2, i, j + 1, k)] - C3[kidx(g, 2, i, j, k)]) * idx2;
        else
          tmp = (C3[kidx(g, 2, i, j + 1, k)] - C3[kidx(g, 2, i, j, k)]) * idx2;
        long f = fidx(g, 0, i, j, k);
        g->bf[0][f] = a1 * g->b0f[0][f] + a2 * g->bf[0][f] + a3 * tmp;
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; k++)
    for (int j = 0; j <= n2; j++)
      for (int i = 0; i < n1; i++) {
        double tmp;
        if (sd == 3)
          tmp = (C1[kidx(g, 0, i, j, k + 1)] - C1[kidx(g, 0, i, j, k)]) * idx3 +
                (C3[kidx(g, 2, i, j, k)] - C3[kidx(g, 2, i + 1, j, k)]) * idx1;
        else
          tmp = (C3[kidx(g, 2, i, j, k)] - C3[kidx(g, 2, i + 1, j, k)]) * idx1;
        long f = fidx(g, 1, i, j, k);
        g->bf[1][f] = a1 * g->b0f[1][f] + a2 * g->bf[1][f] + a3 * tmp;
      }
  if (sd == 3) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          double tmp = (C2[kidx(g, 1, i + 1, j, k)] - C2[kidx(g, 1, i, j, k)]) * idx1 -
                       (C1[kidx(g, 0, i, j + 1, k)] - C1[kidx(g, 0, i, j, k)]) * idx2;
          long f = fidx(g, 2, i, j, k);
          g->bf[2][f] = a1 * g->b0f[2][f] + a2 * g->bf[2][f] + a3 * tmp;
        }
  }
}

/* b_cc = face average (source.F90:12439-12489, :6012-6062) */
static void bcc_from_faces(orc_grid *g) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k);
        g->bcc[c * 3 + 0] = 0.5 * (g->bf[0][fidx(g, 0, i + 1, j, k)] + g->bf[0][fidx(g, 0, i, j, k)]);
        g->bcc[c * 3 + 1] = 0.5 * (g->bf[1][fidx(g, 1, i, j + 1, k)] + g->bf[1][fidx(g, 1, i, j, k)]);
        if (g->sd == 3)
          g->bcc[c * 3 + 2] = 0.5 * (g->bf[2][fidx(g, 2, i, j, k + 1)] + g->bf[2][fidx(g, 2, i, j, k)]);
      }
}

/* ------------------------------------------------------------- the step */

int orc_hydro_step(orc_grid *g) {
  const int nv = g->nv, sd = g->sd, mhd = g->mhd;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const int ip = g->ip;
  double rk[3][3];
  int stages = g->c.rk_stages;
  /* source.F90:7018-7048 */
  if (stages == 2) {
    rk[0][0] = 1.0; rk[0][1] = 0.0; rk[0][2] = -1.0;
    rk[1][0] = 0.5; rk[1][1] = 0.5; rk[1][2] = -0.5;
  } else {
    stages = 3;
    rk[0][0] = 1.0; rk[0][1] = 0.0; rk[0][2] = -1.0;
    rk[1][0] = 0.75; rk[1][1] = 0.25; rk[1][2] = -0.25;
    rk[2][0] = 1.0 / 3.0; rk[2][1] = 2.0 / 3.0; rk[2][2] = -(2.0 / 3.0);
  }
  /* state0 rebuilt from prim + eint (source.F90:7057-7107) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        const double *p = g->prim + c * nv;
        double *s0 = g->state0 + q * nv;
        double rho = p[0], vx1 = p[1], vx2 = p[2], vx3 = (sd == 3) ? p[3] : 0.0;
        s0[0] = rho;
        s0[1] = rho * vx1;
        s0[2] = rho * vx2;
        if (sd == 3) s0[3] = rho * vx3;
        double rhoe = g->eint[q] + 0.5 * rho * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
        if (mhd) {
          const double *b = g->bcc + c * 3;
          double bx3 = (sd == 3) ? b[2] : 0.0;
          rhoe = rhoe + 0.5 * (b[0] * b[0] + b[1] * b[1] + bx3 * bx3);
        }
        if (g->c.evolve_etot) rhoe = rhoe + rho * g->phi_cc[j];
        s0[ip] = rhoe;
        for (int v = g->ias1; v < nv; v++) s0[v] = rho * p[v];
        for (int v = 0; v < nv; v++) g->state[q * nv + v] = s0[v];
      }
  if (mhd)
    for (int d = 0; d < sd; d++) {
      long nf = (long)(n1 + (d == 0)) * (n2 + (d == 1)) * (n3 + (d == 2));
      memcpy(g->b0f[d], g->bf[d], nf * sizeof(double));
    }

  for (int irk = 0; irk < stages; irk++) {
    if (g->c.use_wb) {
      /* source.F90:7151-7168 */
#pragma omp parallel for schedule(static)
      for (int k = 0; k < n3; k++)
        for (int j = 0; j < n2; j++)
          for (int i = 0; i < n1; i++) {
            long c = cidx(g, i, j, k);
            double *p = g->prim + c * nv;
            p[0] = p[0] - g->eq_rho_cc[j];
            p[ip] = p[ip] - g->eq_p_cc[j];
            if (g->c.fasteos) g->gamf[c * 2] = g->gamf[c * 2] - g->eq_ge_cc[j];
          }
    }
    halo_periodic(g, g->prim, nv);
    if (mhd) halo_periodic(g, g->bcc, 3);
    if (TABULATED(g)) halo_periodic(g, g->temp, 1);
    if (g->c.fasteos) halo_periodic(g, g->gamf, 2);
    physical_bc(g);
    if (g->c.shock_flattening) shock_flags(g);
    for (int d = 0; d < sd; d++) sweep(g, d);
    double a1 = rk[irk][0], a2 = rk[irk][1], a3 = rk[irk][2] * g->dt;
    if (mhd) {
      emf_halo(g);
      ct_update(g, a1, a2, a3);
      bcc_from_faces(g);
    }
    /* gravity / heating sources and RK combine, cons -> prim */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
          double *R = g->res + q * nv, *S = g->state + q * nv, *S0 = g->state0 + q * nv;
          double *p = g->prim + c * nv;
          if (g->c.use_gravity) {
            /* source.F90:12606-12644, only g_x2 is non-zero (plane-parallel) */
            double rho = p[0], gx2 = g->grav[j];
            R[2] = R[2] - rho * gx2;
            if (!g->c.evolve_etot) R[ip] = R[ip] - S[2] * gx2;
          }
          if (g->c.use_edot) R[ip] = R[ip] - g->edot[j];
          for (int v = 0; v < nv; v++) S[v] = a1 * S0[v] + a2 * S[v] + a3 * R[v];
          /* source.F90:13533-13686 */
          double rho = S[0], m1 = S[1], m2 = S[2], m3 = (sd == 3) ? S[3] : 0.0;
          double rhoe = S[ip];
          double inv_rho = 1.0 / rho;
          double eint = rhoe - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * inv_rho;
          if (mhd) {
            const double *b = g->bcc + c * 3;
            double bx3 = (sd == 3) ? b[2] : 0.0;
            eint = eint - 0.5 * (b[0] * b[0] + b[1] * b[1] + bx3 * bx3);
          }
          if (g->c.evolve_etot) eint = eint - rho * g->phi_cc[j];
          p[0] = rho;
          p[1] = m1 * inv_rho;
          p[2] = m2 * inv_rho;
          if (sd == 3) p[3] = m3 * inv_rho;
          if (TABULATED(g)) {
            /* ADVECT_YE_IABAR + HELMHOLTZ_EOS (:13573-13578); USE_FASTEOS
             * keeps gamma_e / gamma_c (:13648-13653), without it the final
             * evaluation is pressure only (:13654-13655) */
            double ye = nv > g->ias1 + 1 ? S[g->ias1] * inv_rho : 1.0;
            double abar = nv > g->ias1 + 1 ? rho / S[g->ias1 + 1] : 1.0;
            double zbar = ye * abar;
            double T = g->temp[c], pr, sound;
            orc_helm_rhoe_given(g->helm, rho, eint * inv_rho, abar, zbar, &T, &pr, &sound,
                                g->c.fasteos);
            if (g->c.fasteos) {
              g->gamf[c * 2] = pr / eint + 1.0;
              g->gamf[c * 2 + 1] = sound * sound * rho / pr;
            }
            p[ip] = pr;
            g->temp[c] = T;
          } else {
            double pr = g->gmm1 * eint;
            p[ip] = pr;
            g->temp[c] = pr / (rho * CONST_RGAS * g->inv_mu);
          }
          g->eint[q] = eint;
          for (int v = g->ias1; v < nv; v++) p[v] = S[v] * inv_rho;
        }
  }
  return 0;
}

/* compute_hyperbolic_dt (source.F90:13707-13960) */
int orc_compute_dt(orc_grid *g) {
  const int nv = g->nv, sd = g->sd;
  double smax = 0.0;
#pragma omp parallel for reduction(max : smax) schedule(static)
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k);
        const double *p = g->prim + c * nv;
        double rho = p[0], vx1 = p[1], vx2 = p[2], vx3 = (sd == 3) ? p[3] : 0.0;
        double pr = p[g->ip];
        double cs;
        if (TABULATED(g)) {
          /* :13801-13847: full EoS at the carried T */
          double ye = nv > g->ias1 + 1 ? p[g->ias1] : 1.0;
          double abar = nv > g->ias1 + 1 ? 1.0 / p[g->ias1 + 1] : 1.0;
          double zbar = ye * abar, ei, cv;
          orc_helm_full(g->helm, rho, g->temp[c], abar, zbar, &pr, &ei, &cs, &cv);
        } else {
          cs = sqrt(g->gm * pr / rho);
        }
        if (g->mhd) {
          const double *b = g->bcc + c * 3;
          double bx3 = (sd == 3) ? b[2] : 0.0;
          double ca2 = b[0] * b[0] + b[1] * b[1] + bx3 * bx3;
          cs = sqrt(cs * cs + ca2 / rho);
        }
        double sx1 = (fabs(vx1) + cs) * g->inv_dx[0];
        double sx2 = (fabs(vx2) + cs) * g->inv_dx[1];
        double s;
        if (sd == 3) {
          double sx3 = (fabs(vx3) + cs) * g->inv_dx[2];
          s = sx3 + sx1 + sx2;
        } else {
          s = sx1 + sx2;
        }
        if (s > smax) smax = s;
      }
  g->smax = smax;
  g->dt = g->c.cfl / smax;
  return 0;
}

int orc_init(orc_grid *g) {
  const int nv = g->nv;
  if (g->mhd) bcc_from_faces(g);
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        double rho = g->prim[c * nv], p = g->prim[c * nv + g->ip];
        if (TABULATED(g)) {
          /* step-0 Helmholtz init (source.F90:6279-6285; FASTEOS
           * :6351-6357, otherwise energy only :6358-6361) */
          double abar = nv > g->ias1 + 1 ? 1.0 / g->prim[c * nv + g->ias1 + 1] : 1.0;
          double ye = nv > g->ias1 + 1 ? g->prim[c * nv + g->ias1] : 1.0;
          double zbar = abar * ye;
          double T = g->temp[c], eint, sound;
          orc_helm_rhoP_given(g->helm, rho, p, abar, zbar, &T, &eint, &sound, g->c.fasteos);
          eint = rho * eint;
          if (g->c.fasteos) {
            g->gamf[c * 2] = p / eint + 1.0;
            g->gamf[c * 2 + 1] = sound * sound * rho / p;
          }
          g->eint[q] = eint;
          g->temp[c] = T;
        } else {
          /* ideal gas branch of the step-0 EoS init (source.F90:6373-6376) */
          g->temp[c] = p / (CONST_RGAS * rho * g->inv_mu);
          g->eint[q] = p * g->igmm1;
        }
      }
  g->time = 0.0;
  g->step = 0;
  /* without EVALUATE_PARABOLIC_TIMESTEP dt_par is computed once, at step 0
   * (:6504-6512) */
  if (g->c.thermal_diffusion && g->kappa && !g->c.eval_parabolic_dt) orc_parabolic_dt(g);
  return orc_compute_dt(g);
}

/* Resume from a restart file: read_restart (source.F90:2761-2897) restores
 * prim, the face fields, temp and time, dt, smax, step; time_loop's step /= 0
 * branch (:6413-6500) then rebuilds eint (and the FASTEOS gammas) from
 * (rho, T) by a full Helmholtz evaluation (:6476-6483; the EoS pressure only
 * feeds the gammas) or eint = p / (gamma - 1) (:6492); dt is the file's. */
int orc_resume(orc_grid *g, double time, double dt, double smax, long step) {
  const int nv = g->nv;
  if (g->mhd) bcc_from_faces(g);
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        double rho = g->prim[c * nv], p = g->prim[c * nv + g->ip];
        double eint;
        if (TABULATED(g)) {
          double abar = nv > g->ias1 + 1 ? 1.0 / g->prim[c * nv + g->ias1 + 1] : 1.0;
          double ye = nv > g->ias1 + 1 ? g->prim[c * nv + g->ias1] : 1.0;
          double zbar = abar * ye;
          double T = g->temp[c], pe, sound, cv;
          orc_helm_full(g->helm, rho, T, abar, zbar, &pe, &eint, &sound, &cv);
          eint = rho * eint;
          if (g->c.fasteos) {
            g->gamf[c * 2] = pe / eint + 1.0;
            g->gamf[c * 2 + 1] = sound * sound * rho / pe;
          }
        } else {
          eint = p * g->igmm1;
        }
        g->eint[q] = eint;
      }
  g->time = time;
  g->dt = dt;
  g->smax = smax;
  g->step = step;
  return 0;
}

/* ------------------------------------------------ thermal diffusion (STS) */

#define CONST_RAD 7.565767381646406e-15 /* source.F90:614 */
#define CONST_C 2.99792458e10           /* source.F90:622 */

static inline long kapidx(const orc_grid *g, int i, int j, int k) {
  const int k0 = (g->sd == 3) ? 1 : 0;
  return ((long)(k + k0) * (g->n[1] + 2) + (j + 1)) * (g->n[0] + 2) + (i + 1);
}

int orc_set_kappa(orc_grid *g, const double *kappa) {
  const long n = (long)(g->n[0] + 2) * (g->n[1] + 2) * ((g->sd == 3) ? g->n[2] + 2 : 1);
  const long ni = (long)g->n[0] * g->n[1] * g->n[2];
  if (!g->kappa) {
    g->kappa = (double *)calloc(n, sizeof(double));
    g->e0 = (double *)calloc(ni, sizeof(double));
    g->ejm2 = (double *)calloc(ni, sizeof(double));
    g->me0 = (double *)calloc(ni, sizeof(double));
    g->me = (double *)calloc(ni, sizeof(double));
    if (!g->kappa || !g->e0 || !g->ejm2 || !g->me0 || !g->me) return -1;
  }
  memcpy(g->kappa, kappa, n * sizeof(double));
  return 0;
}

/* compute_parabolic_dt (source.F90:19804-19947), Cartesian, without
 * USE_TIMMES_KAPPA: D = c / (3 kappa rho), dt_par = 1 / max 2 D sum 1/dx^2. */
double orc_parabolic_dt(orc_grid *g) {
  const int nv = g->nv;
  const double i1 = g->inv_dx[0], i2 = g->inv_dx[1], i3 = (g->sd == 3) ? g->inv_dx[2] : 0.0;
  const double inv_dl2 = i1 * i1 + i2 * i2 + i3 * i3;
  double m = 0.0;
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        const double rho = g->prim[cidx(g, i, j, k) * nv];
        const double drad = (1.0 / 3.0) * CONST_C / (g->kappa[kapidx(g, i, j, k)] * rho);
        const double v = 2.0 * drad * inv_dl2;
        m = (m > v) ? m : v;  /* Fortran max(m, v) */
      }
  g->dt_par = 1.0 / m;
  return g->dt_par;
}

/* bcs_thermal_diffusion (source.F90:20837-21068) without internal
 * boundaries: periodic copies of temp; at a physical wall the density
 * ghost is extrapolated linearly (2 rho(lx) - rho(lx+1)) and the
 * temperature ghost copies the wall cell, or with FIX_TEMPERATURE_AT_* is
 * 2 T_wall - T(lx) (the wall value on the face).  One ghost layer, interior
 * transverse ranges. */
static void diff_bcs(orc_grid *g) {
  const int nv = g->nv;
  halo_periodic(g, g->temp, 1);
  for (int d = 0; d < g->sd; d++) {
    if (g->c.bc[2 * d] == ORC_BC_PERIODIC) continue;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int n = g->n[d];
    for (int side = 0; side < 2; side++) {
      const double tw = g->c.fix_t[2 * d + side];
      for (int b = 0; b < g->n[t2]; b++)
        for (int a = 0; a < g->n[t1]; a++) {
          int id[3], i0[3], i1[3];
          id[t1] = i0[t1] = i1[t1] = a;
          id[t2] = i0[t2] = i1[t2] = b;
          id[d] = side == 0 ? -1 : n;
          i0[d] = side == 0 ? 0 : n - 1;
          i1[d] = side == 0 ? 1 : n - 2;
          const long cg = cidx(g, id[0], id[1], id[2]);
          const long c0 = cidx(g, i0[0], i0[1], i0[2]), c1 = cidx(g, i1[0], i1[1], i1[2]);
          g->prim[cg * nv] = 2.0 * g->prim[c0 * nv] - g->prim[c1 * nv];
          if (tw == tw)
            g->temp[cg] = 2.0 * tw - g->temp[c0];
          else
            g->temp[cg] = g->temp[c0];
        }
    }
  }
}

/* One face's radiative flux Krad K_int grad T between cells m and p
 * (source.F90:20488-20509): K = T^3 / (rho kappa), harmonic mean. */
static inline double diff_flux(const orc_grid *g, long cm, long cp, long km, long kp,
                               double inv_dl) {
  const double krad = (4.0 / 3.0) * CONST_RAD * CONST_C;
  const double tm = g->temp[cm], tp = g->temp[cp];
  const double kmn = tm * tm * tm / (g->prim[cm * g->nv] * g->kappa[km]);
  const double kpl = tp * tp * tp / (g->prim[cp * g->nv] * g->kappa[kp]);
  const double kint = 2.0 * (kmn * kpl) / (kmn + kpl);
  const double grad = (tp - tm) * inv_dl;
  return krad * kint * grad;
}

/* residuals_thermal_diffusion (source.F90:20394-20835), Cartesian: me =
 * sum_d (fe(+) - fe(-)) / dx_d, x1 first. */
static void diff_residual(orc_grid *g) {
  const int sd = g->sd;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        const long q = iidx(g, i, j, k);
        const long c = cidx(g, i, j, k), kc = kapidx(g, i, j, k);
        double m;
        {
          const double fm = diff_flux(g, cidx(g, i - 1, j, k), c, kapidx(g, i - 1, j, k), kc,
                                      g->inv_dx[0]);
          const double fp = diff_flux(g, c, cidx(g, i + 1, j, k), kc, kapidx(g, i + 1, j, k),
                                      g->inv_dx[0]);
          m = (fp - fm) * g->inv_dx[0];
        }
        {
          const double fm = diff_flux(g, cidx(g, i, j - 1, k), c, kapidx(g, i, j - 1, k), kc,
                                      g->inv_dx[1]);
          const double fp = diff_flux(g, c, cidx(g, i, j + 1, k), kc, kapidx(g, i, j + 1, k),
                                      g->inv_dx[1]);
          m = m + (fp - fm) * g->inv_dx[1];
        }
        if (sd == 3) {
          const double fm = diff_flux(g, cidx(g, i, j, k - 1), c, kapidx(g, i, j, k - 1), kc,
                                      g->inv_dx[2]);
          const double fp = diff_flux(g, c, cidx(g, i, j, k + 1), kc, kapidx(g, i, j, k + 1),
                                      g->inv_dx[2]);
          m = m + (fp - fm) * g->inv_dx[2];
        }
        g->me[q] = m;
      }
}

/* eint_to_temp (source.F90:20271-20392): ideal gas closed form, Helmholtz
 * helm_rhoe_given from the carried T (energy only). */
static void eint_to_temp(orc_grid *g) {
  const int nv = g->nv;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        const long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        const double rho = g->prim[c * nv], eint = g->eint[q];
        if (TABULATED(g)) {
          const double ye = nv > g->ias1 + 1 ? g->prim[c * nv + g->ias1] : 1.0;
          const double abar = nv > g->ias1 + 1 ? 1.0 / g->prim[c * nv + g->ias1 + 1] : 1.0;
          const double zbar = ye * abar;
          double T = g->temp[c], pr, sound;
          orc_helm_rhoe_given(g->helm, rho, eint / rho, abar, zbar, &T, &pr, &sound, 0);
          g->temp[c] = T;
        } else {
          g->temp[c] = g->gmm1 * eint / (g->inv_mu * rho * CONST_RGAS);
        }
      }
}

/* thermal_diffusion_step (source.F90:19949-20269), evolving eint (no
 * STS_EVOLVE_TEMP): RKL2 super-time-stepping over dt/2 with
 * s = max(3, 1 + floor((sqrt(9 + 16 dt/dt_par) - 1) / 2)) stages
 * (Meyer, Balsara & Aslam 2014), then T from eint and p (Helmholtz: eint,
 * p and the FASTEOS gammas) from (rho, T). */
int orc_diffusion_step(orc_grid *g, int strang) {
  (void)strang;
  if (!g->kappa) return -1;
  const int nv = g->nv, ip = g->ip;
  const long ni = (long)g->n[0] * g->n[1] * g->n[2];
  const double fac = g->dt / g->dt_par;
  const double half_dt = 0.5 * g->dt;
  int s = 1 + (int)floor(0.5 * (sqrt(9.0 + 16.0 * fac) - 1.0));
  if (s < 3) s = 3;
  const double w1 = 4.0 / ((double)(s * s + s) - 2.0);
  const double b1 = 1.0 / 3.0, mut1 = b1 * w1;
  for (long q = 0; q < ni; q++) g->e0[q] = g->eint[q];
  halo_periodic(g, g->prim, nv);
  diff_bcs(g);
  diff_residual(g);
  for (long q = 0; q < ni; q++) {
    g->me0[q] = g->me[q];
    const double e1 = g->e0[q] + half_dt * mut1 * g->me0[q];
    g->ejm2[q] = g->e0[q];
    g->eint[q] = e1;
  }
  double bjm2 = 1.0 / 3.0, bjm1 = 1.0 / 3.0;
  for (int j = 2; j <= s; j++) {
    const double bj = ((double)(j * j + j) - 2.0) / (2.0 * j * (j + 1.0));
    const double muj = ((2.0 * j - 1.0) / j) * bj / bjm1;
    const double mutj = muj * w1;
    const double nuj = -(((j - 1.0) / j) * bj / bjm2);
    const double ajm1 = 1.0 - bjm1;
    const double gammatj = -ajm1 * mutj;
    eint_to_temp(g);
    diff_bcs(g);
    diff_residual(g);
    for (long q = 0; q < ni; q++) {
      const double ejm1 = g->eint[q];
      g->eint[q] = muj * ejm1 + nuj * g->ejm2[q] + (1.0 - muj - nuj) * g->e0[q] +
                   half_dt * mutj * g->me[q] + half_dt * gammatj * g->me0[q];
      g->ejm2[q] = ejm1;
    }
    bjm2 = bjm1;
    bjm1 = bj;
  }
  eint_to_temp(g);
  for (int k = 0; k < g->n[2]; k++)
    for (int j = 0; j < g->n[1]; j++)
      for (int i = 0; i < g->n[0]; i++) {
        const long c = cidx(g, i, j, k), q = iidx(g, i, j, k);
        const double rho = g->prim[c * nv], T = g->temp[c];
        if (TABULATED(g)) {
          const double ye = nv > g->ias1 + 1 ? g->prim[c * nv + g->ias1] : 1.0;
          const double abar = nv > g->ias1 + 1 ? 1.0 / g->prim[c * nv + g->ias1 + 1] : 1.0;
          const double zbar = abar * ye;
          double pr, e, sound, cv;
          orc_helm_full(g->helm, rho, T, abar, zbar, &pr, &e, &sound, &cv);
          const double eint = rho * e;
          g->prim[c * nv + ip] = pr;
          g->eint[q] = eint;
          if (g->c.fasteos) {
            g->gamf[c * 2] = pr / eint + 1.0;
            g->gamf[c * 2 + 1] = sound * sound * rho / pr;
          }
        } else {
          g->prim[c * nv + ip] = CONST_RGAS * rho * T * g->inv_mu;
        }
      }
  return 0;
}

int orc_run(orc_grid *g, int nsteps) {
  const int diff = g->c.thermal_diffusion && g->kappa;
  for (int s = 0; s < nsteps; s++) {
    if (!(g->time < g->c.tmax)) break;                             /* :6525 */
    if (g->time + g->dt > g->c.tmax) g->dt = g->c.tmax - g->time; /* :6662-6664 */
    if (diff) {                                                    /* :6672-6679 */
      if (g->c.eval_parabolic_dt) orc_parabolic_dt(g);
      orc_diffusion_step(g, 1);
    }
    orc_hydro_step(g);
    if (diff) orc_diffusion_step(g, 2);                             /* :6685-6687 */
    g->time = g->time + g->dt;
    orc_compute_dt(g);
    g->step++;
  }
  return 0;
}
