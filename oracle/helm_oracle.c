/*
 * helm_oracle.c -- CPU FP64 restatement of the reference's Helmholtz EoS
 * (source/source.F90:23074-25058).
 *
 * TEST INFRASTRUCTURE ONLY (checker of the HIP EoS and of the Helmholtz
 * hydro path).  The table is whatever the caller loads (the build generates
 * it with phlegethon_amd/helmgen); this file only restates how the
 * reference interpolates and inverts it.
 *
 * Table storage follows the reference's internal variable ids
 * (source/source.F90:525-545): f block  1 f, 2 dfdT, 3 d2fdT2, 4 dfdrho,
 * 5 d2fdrho2, 6 d2fdrhodT, 7 d3fdrho2dT, 8 d3fdrhodT2, 9 d4fdrho2dT2;
 * dP/drho block 1 c, 2 dcdT, 3 dcdrho, 4 d2cdrhodT.
 *
 * Parity pins (tests/test_helmholtz_oracle.py): closed forms of the e-/e+
 * gas (nondegenerate relativistic Maxwell-Juettner, T = 0 Chandrasekhar,
 * ultra-relativistic pair plasma), exactness on a table generated from a
 * biquintic polynomial, interpolation between nodes against direct table
 * values at the cell centres, Newton round trips and the reset path.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RGAS 8.31446261815324e7   /* source.F90:605-616 */
#define CRAD 7.565767381646406e-15
#define CLIGHT 2.99792458e10
#define CLIGHT2 8.987551787368177e20

struct orc_helm {
  int nr, nt;
  double ldlo, ldhi, ltlo, lthi, dr, dt;
  double *rho, *T;   /* node values */
  double *f;         /* [j][i][9] internal ids */
  double *c;         /* [j][i][4] internal ids */
  double tl, th;     /* Tl_eos, Th_eos (source.F90:750-751) */
  double lrho1, lT1; /* log10 of the first nodes */
  int pig;           /* PIG_EOS table: the whole gas in f(rho, T), Ye = 1, no ions */
};

/* file order: f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt -> internal */
static const int F_FILE_TO_ID[9] = {0, 3, 1, 4, 2, 5, 6, 7, 8};
/* file order: c, cd, ct, cdt -> internal (c, ct, cd, cdt) */
static const int C_FILE_TO_ID[4] = {0, 2, 1, 3};

orc_helm *orc_helm_create(int nr, int nt, double ldlo, double ldhi, double ltlo,
                          double lthi, const double *f9_file, const double *c4_file) {
  orc_helm *h = (orc_helm *)calloc(1, sizeof(orc_helm));
  h->nr = nr;
  h->nt = nt;
  h->ldlo = ldlo;
  h->ldhi = ldhi;
  h->ltlo = ltlo;
  h->lthi = lthi;
  /* source.F90:23094-23104 */
  h->dt = (lthi - ltlo) / (double)(nt - 1);
  h->dr = (ldhi - ldlo) / (double)(nr - 1);
  h->rho = (double *)malloc(sizeof(double) * nr);
  h->T = (double *)malloc(sizeof(double) * nt);
  for (int j = 0; j < nt; j++) h->T[j] = pow(10.0, ltlo + (double)j * h->dt);
  for (int i = 0; i < nr; i++) h->rho[i] = pow(10.0, ldlo + (double)i * h->dr);
  h->f = (double *)malloc(sizeof(double) * 9 * nr * nt);
  h->c = (double *)malloc(sizeof(double) * 4 * nr * nt);
  for (long p = 0; p < (long)nr * nt; p++) {
    for (int v = 0; v < 9; v++) h->f[9 * p + F_FILE_TO_ID[v]] = f9_file[9 * p + v];
    for (int v = 0; v < 4; v++) h->c[4 * p + C_FILE_TO_ID[v]] = c4_file[4 * p + v];
  }
  h->tl = pow(10.0, ltlo);
  h->th = pow(10.0, lthi);
  h->lrho1 = log10(h->rho[0]);
  h->lT1 = log10(h->T[0]);
  return h;
}

void orc_helm_destroy(orc_helm *h) {
  if (!h) return;
  free(h->rho); free(h->T); free(h->f); free(h->c);
  free(h);
}

double orc_helm_node(const orc_helm *h, int which, int idx) {
  return which == 0 ? h->rho[idx] : h->T[idx];
}

/* cell location (source.F90:23773-23784); the index is kept inside the
 * table (the reference does not clamp; out-of-table inputs are undefined
 * there). */
typedef struct {
  int ih, jh;
  double x, y, drho, dT;
} cellpos;

static cellpos locate(const orc_helm *h, double rhos, double T) {
  cellpos c;
  c.ih = (int)((log10(rhos) - h->lrho1) / h->dr);
  c.jh = (int)((log10(T) - h->lT1) / h->dt);
  if (c.ih < 0) c.ih = 0;
  if (c.ih > h->nr - 2) c.ih = h->nr - 2;
  if (c.jh < 0) c.jh = 0;
  if (c.jh > h->nt - 2) c.jh = h->nt - 2;
  double tmp = h->rho[c.ih];
  c.drho = h->rho[c.ih + 1] - tmp;
  c.x = (rhos - tmp) / c.drho;
  tmp = h->T[c.jh];
  c.dT = h->T[c.jh + 1] - tmp;
  c.y = (T - tmp) / c.dT;
  return c;
}

/* quintic Hermite basis and derivatives of one coordinate (source.F90
 * :23832-23862 value, :23900-23906 first, :24030-24036 second derivative) */
typedef struct {
  double p0, p1, p2, p0m, p1m, p2m;
} basis;

static void quintic(double x, double d, basis *v, basis *dv, basis *ddv) {
  const double id = 1.0 / d, id2 = id * id, d2 = d * d;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
  const double o = 1.0 - x;
  const double o2 = o * o, o3 = o2 * o, o4 = o3 * o, o5 = o4 * o;
  v->p0 = -6.0 * x5 + 15.0 * x4 - 10.0 * x3 + 1.0;
  v->p1 = (-3.0 * x5 + 8.0 * x4 - 6.0 * x3 + x) * d;
  v->p2 = (0.5 * (-x5 + 3.0 * x4 - 3.0 * x3 + x2)) * d2;
  v->p0m = -6.0 * o5 + 15.0 * o4 - 10.0 * o3 + 1.0;
  v->p1m = (3.0 * o5 - 8.0 * o4 + 6.0 * o3 - o) * d;
  v->p2m = (0.5 * (-o5 + 3.0 * o4 - 3.0 * o3 + o2)) * d2;
  if (dv) {
    dv->p0 = (-30.0 * x4 + 60.0 * x3 - 30.0 * x2) * id;
    dv->p1 = (-15.0 * x4 + 32.0 * x3 - 18.0 * x2 + 1.0);
    dv->p2 = (0.5 * (-5.0 * x4 + 12.0 * x3 - 9.0 * x2 + 2.0 * x)) * d;
    dv->p0m = -(-30.0 * o4 + 60.0 * o3 - 30.0 * o2) * id;
    dv->p1m = (-15.0 * o4 + 32.0 * o3 - 18.0 * o2 + 1.0);
    dv->p2m = -(0.5 * (-5.0 * o4 + 12.0 * o3 - 9.0 * o2 + 2.0 * o)) * d;
  }
  if (ddv) {
    ddv->p0 = (-120.0 * x3 + 180.0 * x2 - 60.0 * x) * id2;
    ddv->p1 = (-60.0 * x3 + 96.0 * x2 - 36.0 * x) * id;
    ddv->p2 = (0.5 * (-20.0 * x3 + 36.0 * x2 - 18.0 * x + 2.0));
    ddv->p0m = (-120.0 * o3 + 180.0 * o2 - 60.0 * o) * id2;
    ddv->p1m = -(-60.0 * o3 + 96.0 * o2 - 36.0 * o) * id;
    ddv->p2m = (0.5 * (-20.0 * o3 + 36.0 * o2 - 18.0 * o + 2.0));
  }
}

/* the 36-term biquintic sum in the reference's term order: variable v
 * pairs the rho basis of order RX[v] with the T basis of order TY[v]; for
 * each v the corners (ih,jh) (ih+1,jh) (ih,jh+1) (ih+1,jh+1). */
static const int RX[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2};
static const int TY[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};

static double sel(const basis *b, int order, int minus) {
  if (order == 0) return minus ? b->p0m : b->p0;
  if (order == 1) return minus ? b->p1m : b->p1;
  return minus ? b->p2m : b->p2;
}

static double sum36(const orc_helm *h, const cellpos *c, const basis *r, const basis *t) {
  const double *n00 = h->f + 9 * ((long)c->jh * h->nr + c->ih);
  const double *n10 = n00 + 9;
  const double *n01 = n00 + 9 * h->nr;
  const double *n11 = n01 + 9;
  double s = 0.0;
  for (int v = 0; v < 9; v++) {
    const double rb = sel(r, RX[v], 0), rbm = sel(r, RX[v], 1);
    const double tb = sel(t, TY[v], 0), tbm = sel(t, TY[v], 1);
    const double t0 = n00[v] * rb * tb, t1 = n10[v] * rbm * tb;
    const double t2 = n01[v] * rb * tbm, t3 = n11[v] * rbm * tbm;
    if (v == 0) s = t0; else s = s + t0;
    s = s + t1;
    s = s + t2;
    s = s + t3;
  }
  return s;
}

/* bicubic Hermite of the dP/drho block (source.F90:24078-24125) */
static double sum16(const orc_helm *h, const cellpos *c) {
  const double x = c->x, y = c->y, o = 1.0 - x, oy = 1.0 - y;
  const double x2 = x * x, x3 = x2 * x, o2 = o * o, o3 = o2 * o;
  const double y2 = y * y, y3 = y2 * y, oy2 = oy * oy, oy3 = oy2 * oy;
  const double p0r = 2.0 * x3 - 3.0 * x2 + 1.0;
  const double p1r = (x3 - 2.0 * x2 + x) * c->drho;
  const double p0mr = 2.0 * o3 - 3.0 * o2 + 1.0;
  const double p1mr = -(o3 - 2.0 * o2 + o) * c->drho;
  const double p0t = 2.0 * y3 - 3.0 * y2 + 1.0;
  const double p1t = (y3 - 2.0 * y2 + y) * c->dT;
  const double p0mt = 2.0 * oy3 - 3.0 * oy2 + 1.0;
  const double p1mt = -(oy3 - 2.0 * oy2 + oy) * c->dT;
  const double *n00 = h->c + 4 * ((long)c->jh * h->nr + c->ih);
  const double *n10 = n00 + 4, *n01 = n00 + 4 * h->nr, *n11 = n01 + 4;
  const double rb[4] = {p0r, p0r, p1r, p1r}, rbm[4] = {p0mr, p0mr, p1mr, p1mr};
  const double tb[4] = {p0t, p1t, p0t, p1t}, tbm[4] = {p0mt, p1mt, p0mt, p1mt};
  double s = 0.0;
  for (int v = 0; v < 4; v++) {
    const double t0 = n00[v] * rb[v] * tb[v], t1 = n10[v] * rbm[v] * tb[v];
    const double t2 = n01[v] * rb[v] * tbm[v], t3 = n11[v] * rbm[v] * tbm[v];
    if (v == 0) s = t0; else s = s + t0;
    s = s + t1;
    s = s + t2;
    s = s + t3;
  }
  return s;
}

/* ---------------------------------------------------------------- PIG EoS */

/* pig_rhoT_given_full (source.F90:25457-25900): the table holds the free
 * energy of the whole partially ionised gas at rho (Ye = 1), radiation is
 * added; no ion terms. */
static void pig_full(const orc_helm *h, double rho, double T, double *P, double *E,
                     double *sound, double *cv) {
  const double rhos = rho;
  cellpos c = locate(h, rhos, T);
  basis r, dr, t, dt, ddt;
  quintic(c.x, c.drho, &r, &dr, NULL);
  quintic(c.y, c.dT, &t, &dt, &ddt);
  const double f = sum36(h, &c, &r, &t);
  const double df_drho = sum36(h, &c, &dr, &t);
  const double df_dT = sum36(h, &c, &r, &dt);
  const double df_drhodT = sum36(h, &c, &dr, &dt);
  const double df_dT2 = sum36(h, &c, &r, &ddt);
  const double dP_drho = sum16(h, &c);
  const double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double s = -df_dT;
  const double ds_dT = -df_dT2;
  double e = f + T * s;
  double dE_dT = T * ds_dT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  p = p + CRAD * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + (4.0 / 3.0) * CRAD * T3;
  e = e + CRAD * T4 / rho;
  dE_dT = dE_dT + 4.0 * CRAD * T3 / rho;
  const double zz = p / rho;
  const double chiT = T / p * dP_dT;
  const double chirho = dP_drho / zz;
  const double x = zz * chiT / (T * dE_dT);
  const double gam1 = chiT * x + chirho;
  const double z = 1.0 + (e + CLIGHT2) / zz;
  *P = p;
  *E = e;
  *cv = dE_dT;
  *sound = CLIGHT * sqrt(gam1 / z);
}

/* pig_rhoT_given1 (source.F90:26094-26353): P and dP/dT */
static void pig_given1(const orc_helm *h, double rho, double T, double *P, double *dPdT,
                       int ret) {
  const double rhos = rho;
  cellpos c = locate(h, rhos, T);
  basis r, dr, t, dt;
  quintic(c.x, c.drho, &r, &dr, NULL);
  quintic(c.y, c.dT, &t, &dt, NULL);
  const double df_drho = sum36(h, &c, &dr, &t);
  double df_drhodT = 0.0;
  if (ret) df_drhodT = sum36(h, &c, &dr, &dt);
  const double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  p = p + CRAD * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + (4.0 / 3.0) * CRAD * T3;
  *P = p;
  *dPdT = dP_dT;
}

/* pig_rhoT_given2 (source.F90:26355-26653): E and dE/dT */
static void pig_given2(const orc_helm *h, double rho, double T, double *E, double *dEdT,
                       int ret) {
  const double rhos = rho;
  cellpos c = locate(h, rhos, T);
  basis r, t, dt, ddt;
  quintic(c.x, c.drho, &r, NULL, NULL);
  quintic(c.y, c.dT, &t, &dt, &ddt);
  const double f = sum36(h, &c, &r, &t);
  const double df_dT = sum36(h, &c, &r, &dt);
  double df_dT2 = 0.0;
  if (ret) df_dT2 = sum36(h, &c, &r, &ddt);
  const double s = -df_dT;
  const double ds_dT = -df_dT2;
  double e = f + T * s;
  double dE_dT = T * ds_dT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  e = e + CRAD * T4 / rho;
  dE_dT = dE_dT + 4.0 * CRAD * T3 / rho;
  *E = e;
  *dEdT = dE_dT;
}

void orc_helm_set_pig(orc_helm *h, int pig) { h->pig = pig != 0; }

/* helm_rhoT_given_full (source.F90:23686-24232), no Coulomb corrections */
void orc_helm_full(const orc_helm *h, double rho, double T, double abar, double zbar,
                   double *P, double *E, double *sound, double *cv) {
  if (h->pig) {
    pig_full(h, rho, T, P, E, sound, cv);
    return;
  }
  const double Ye = zbar / abar, rhos = rho * Ye;
  cellpos c = locate(h, rhos, T);
  basis r, dr, t, dt, ddt;
  quintic(c.x, c.drho, &r, &dr, NULL);
  quintic(c.y, c.dT, &t, &dt, &ddt);
  const double f = sum36(h, &c, &r, &t);
  const double df_drho = sum36(h, &c, &dr, &t);
  const double df_dT = sum36(h, &c, &r, &dt);
  const double df_drhodT = sum36(h, &c, &dr, &dt);
  const double df_dT2 = sum36(h, &c, &r, &ddt);
  double dP_drho = sum16(h, &c);
  dP_drho = Ye * dP_drho;
  double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double s = -Ye * df_dT;
  const double ds_dT = -Ye * df_dT2;
  tmp = Ye * Ye;
  const double ds_drho = -df_drhodT * tmp;
  double e = Ye * f + T * s;
  double dE_dT = T * ds_dT;
  double dE_drho = df_drho * tmp + T * ds_drho;
  tmp = 1.0 / abar;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double pion = RGAS * rho * T * tmp;
  const double dpiondt = RGAS * rho * tmp;
  const double dpiondd = RGAS * T * tmp;
  p = p + pion + CRAD * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + dpiondt + (4.0 / 3.0) * CRAD * T3;
  dP_drho = dP_drho + dpiondd;
  e = e + 1.5 * RGAS * T * tmp + CRAD * T4 / rho;
  dE_dT = dE_dT + 1.5 * RGAS * tmp + 4.0 * CRAD * T3 / rho;
  dE_drho = dE_drho - CRAD * T4 / (rho * rho);
  const double zz = p / rho;
  const double chiT = T / p * dP_dT;
  const double chirho = dP_drho / zz;
  const double x = zz * chiT / (T * dE_dT);
  const double gam1 = chiT * x + chirho;
  const double z = 1.0 + (e + CLIGHT2) / zz;
  *P = p;
  *E = e;
  *cv = dE_dT;
  *sound = CLIGHT * sqrt(gam1 / z);
  (void)dE_drho;
}

/* helm_rhoT_given1 (source.F90:24234-24569): P and dP/dT */
void orc_helm_given1(const orc_helm *h, double rho, double T, double abar, double zbar,
                     double *P, double *dPdT, int ret) {
  if (h->pig) {
    pig_given1(h, rho, T, P, dPdT, ret);
    return;
  }
  const double Ye = zbar / abar, rhos = rho * Ye;
  cellpos c = locate(h, rhos, T);
  basis r, dr, t, dt;
  quintic(c.x, c.drho, &r, &dr, NULL);
  quintic(c.y, c.dT, &t, &dt, NULL);
  const double df_drho = sum36(h, &c, &dr, &t);
  double df_drhodT = 0.0;
  if (ret) df_drhodT = sum36(h, &c, &dr, &dt);
  double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  tmp = 1.0 / abar;
  const double pion = RGAS * rho * T * tmp;
  const double dpiondt = RGAS * rho * tmp;
  p = p + pion + CRAD * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + dpiondt + (4.0 / 3.0) * CRAD * T3;
  *P = p;
  *dPdT = dP_dT;
}

/* helm_rhoT_given2 (source.F90:24571-24941): E and dE/dT */
void orc_helm_given2(const orc_helm *h, double rho, double T, double abar, double zbar,
                     double *E, double *dEdT, int ret) {
  if (h->pig) {
    pig_given2(h, rho, T, E, dEdT, ret);
    return;
  }
  const double Ye = zbar / abar, rhos = rho * Ye;
  cellpos c = locate(h, rhos, T);
  basis r, t, dt, ddt;
  quintic(c.x, c.drho, &r, NULL, NULL);
  quintic(c.y, c.dT, &t, &dt, &ddt);
  const double f = sum36(h, &c, &r, &t);
  const double df_dT = sum36(h, &c, &r, &dt);
  double df_dT2 = 0.0;
  if (ret) df_dT2 = sum36(h, &c, &r, &ddt);
  const double s = -Ye * df_dT;
  const double ds_dT = -Ye * df_dT2;
  double e = Ye * f + T * s;
  double dE_dT = T * ds_dT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double tmp = 1.0 / abar;
  e = e + 1.5 * RGAS * T * tmp + CRAD * T4 / rho;
  dE_dT = dE_dT + 1.5 * RGAS * tmp + 4.0 * CRAD * T3 / rho;
  *E = e;
  *dEdT = dE_dT;
}

/* helm_rhoe_given (source.F90:24982-25019): Newton T(rho, e), warm start T;
 * returns the number of iterations taken. */
int orc_helm_rhoe_given(const orc_helm *h, double rho, double E0, double abar, double zbar,
                        double *T, double *P, double *sound, int ret) {
  double e, dEdT, cv, dPdT;
  const double Told = *T;
  int iter;
  *sound = 0.0;
  for (iter = 1; iter <= 100; iter++) {
    orc_helm_given2(h, rho, *T, abar, zbar, &e, &dEdT, 1);
    const double res = e - E0;
    *T = *T - res / dEdT;
    double err = fabs(res / E0);
    if ((*T < h->tl) || (*T > h->th)) {
      *T = Told;
      err = 1e-15;
    }
    if (err < 1e-11) break;
  }
  if (ret)
    orc_helm_full(h, rho, *T, abar, zbar, P, &e, sound, &cv);
  else
    orc_helm_given1(h, rho, *T, abar, zbar, P, &dPdT, 0);
  return iter > 100 ? 100 : iter;
}

/* helm_rhoP_given (source.F90:24943-24980): Newton T(rho, P) */
int orc_helm_rhoP_given(const orc_helm *h, double rho, double P0, double abar, double zbar,
                        double *T, double *E, double *sound, int ret) {
  double p, dPdT, cv, dEdT;
  const double Told = *T;
  int iter;
  *sound = 0.0;
  for (iter = 1; iter <= 100; iter++) {
    orc_helm_given1(h, rho, *T, abar, zbar, &p, &dPdT, 1);
    const double res = p - P0;
    *T = *T - res / dPdT;
    double err = fabs(res / P0);
    if ((*T < h->tl) || (*T > h->th)) {
      *T = Told;
      err = 1e-15;
    }
    if (err < 1e-11) break;
  }
  if (ret)
    orc_helm_full(h, rho, *T, abar, zbar, &p, E, sound, &cv);
  else
    orc_helm_given2(h, rho, *T, abar, zbar, E, &dEdT, 0);
  return iter > 100 ? 100 : iter;
}

/* batched helpers for the tests */
void orc_helm_full_batch(const orc_helm *h, int n, const double *rho, const double *T,
                         const double *abar, const double *zbar, double *P, double *E,
                         double *cs, double *cv) {
  for (int i = 0; i < n; i++) orc_helm_full(h, rho[i], T[i], abar[i], zbar[i], P + i, E + i, cs + i, cv + i);
}

void orc_helm_rhoe_batch(const orc_helm *h, int n, const double *rho, const double *e,
                         const double *abar, const double *zbar, double *T, double *P,
                         double *cs, int ret, int *iters) {
  for (int i = 0; i < n; i++) {
    int it = orc_helm_rhoe_given(h, rho[i], e[i], abar[i], zbar[i], T + i, P + i, cs + i, ret);
    if (iters) iters[i] = it;
  }
}

void orc_helm_rhoP_batch(const orc_helm *h, int n, const double *rho, const double *p,
                         const double *abar, const double *zbar, double *T, double *E,
                         double *cs, int ret, int *iters) {
  for (int i = 0; i < n; i++) {
    int it = orc_helm_rhoP_given(h, rho[i], p[i], abar[i], zbar[i], T + i, E + i, cs + i, ret);
    if (iters) iters[i] = it;
  }
}

/* raw interpolated table sums at (rhoYe, T): f, df/drho, df/dT, d2f/drhodT,
 * d2f/dT2 and the bicubic dP/d(rhoYe) (tests only). */
void orc_helm_interp(const orc_helm *h, double rhos, double T, double *out6) {
  cellpos c = locate(h, rhos, T);
  basis r, dr, t, dt, ddt;
  quintic(c.x, c.drho, &r, &dr, NULL);
  quintic(c.y, c.dT, &t, &dt, &ddt);
  out6[0] = sum36(h, &c, &r, &t);
  out6[1] = sum36(h, &c, &dr, &t);
  out6[2] = sum36(h, &c, &r, &dt);
  out6[3] = sum36(h, &c, &dr, &dt);
  out6[4] = sum36(h, &c, &r, &ddt);
  out6[5] = sum16(h, &c);
}
