/*
 * helmgen.c -- generator of the Helmholtz free-energy table of the
 * electron-positron gas used by the Helmholtz EoS (the reference loads it
 * from helm_table.dat, source/source.F90:23077-23156; that file is a
 * download and absent here, so it is computed from first principles).
 *
 * Physics (Timmes & Arnett 1999; Timmes & Swesty 2000, ApJS 126, 501):
 * non-interacting electrons and positrons of arbitrary degeneracy and
 * relativity in pair equilibrium.  With beta = kT/(m c^2) and the
 * degeneracy parameter eta (chemical potential without rest mass, in kT),
 *   n   = Kn beta^{3/2} [F_{1/2}(eta,beta) + beta F_{3/2}(eta,beta)]
 *   P   = (2/3) Kn m c^2 beta^{5/2} [F_{3/2} + beta/2 F_{5/2}]
 *   Kn  = 8 pi sqrt(2) (m c / h)^3
 *   F_k(eta,beta) = int_0^inf x^k sqrt(1 + beta x/2) / (exp(x - eta) + 1) dx
 * positrons take eta_p = -eta - 2/beta.  Charge neutrality fixes eta:
 *   n_e - n_p = rhoYe * N_A.
 * With the energy zero that counts 2 m c^2 per pair, mu_e = eta kT and
 * mu_p = -mu_e, so the free energy per unit volume is F = eta kT (n_e - n_p)
 * - P, and the table holds f = F / (rho Ye) as a function of (rho Ye, T):
 *   f = eta k T N_A - P / (rho Ye).
 * The EoS then forms P = (rho Ye)^2 df/d(rho Ye), s = -Ye df/dT,
 * E = Ye f + T s.
 *
 * Derivatives: every quantity is carried as a bivariate truncated Taylor
 * polynomial in (d rhoYe, d T) of degree <= 3 in rhoYe and <= 2 in T, so the
 * nine table entries (f and its derivatives up to d4f/drho2dT2) and the
 * four dP/d(rhoYe) entries (up to d2/drho dT) come out of one evaluation.
 * eta(rhoYe, T) is obtained by a scalar Newton solve followed by Newton
 * iterations in the Taylor algebra (each adds one order).  The
 * Fermi-Dirac integrals are evaluated by composite Gauss-Legendre panels
 * graded around the Fermi surface plus a Gauss-Laguerre tail, with the
 * Taylor-valued integrand.
 *
 * Output: the reference's ASCII layout (9-column f block, j (T) outer, i
 * (rho) inner; then the 4-column dP/drho block; then two 4-column blocks of
 * eta and number density for format completeness) and a binary twin
 * [header][f9][dpdf4] read by the library.
 *
 * usage: helmgen nrho nT ldlo ldhi ltlo lthi out.bin [out.dat]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants: the reference's values where it has one (source/source.F90:604-616) */
#define C_LIGHT 2.99792458e10
#define K_B 1.380650424e-16
#define H_PL 6.62606896e-27
#define N_AV 6.02214076e23
#define M_ELEC 9.10938215e-28
#define PI_ 3.141592653589793238

#define DR 3 /* max degree in rhoYe */
#define DT 2 /* max degree in T */
typedef struct {
  double c[DR + 1][DT + 1];
} tay;

static tay tconst(double v) {
  tay r;
  memset(&r, 0, sizeof r);
  r.c[0][0] = v;
  return r;
}
static tay tadd(tay a, tay b) {
  for (int i = 0; i <= DR; i++)
    for (int j = 0; j <= DT; j++) a.c[i][j] += b.c[i][j];
  return a;
}
static tay tsub(tay a, tay b) {
  for (int i = 0; i <= DR; i++)
    for (int j = 0; j <= DT; j++) a.c[i][j] -= b.c[i][j];
  return a;
}
static tay tscale(tay a, double s) {
  for (int i = 0; i <= DR; i++)
    for (int j = 0; j <= DT; j++) a.c[i][j] *= s;
  return a;
}
static tay tmul(const tay *a, const tay *b) {
  tay r;
  memset(&r, 0, sizeof r);
  for (int i = 0; i <= DR; i++)
    for (int j = 0; j <= DT; j++) {
      double s = 0.0;
      for (int p = 0; p <= i; p++)
        for (int q = 0; q <= j; q++) s += a->c[p][q] * b->c[i - p][j - q];
      r.c[i][j] = s;
    }
  return r;
}
/* series of g(a) = sum_n coef[n] (a - a0)^n, coef up to order DR+DT */
static tay tcompose(const tay *a, const double *coef) {
  tay z = *a;
  z.c[0][0] = 0.0;
  tay r = tconst(coef[0]);
  tay zp = tconst(1.0);
  for (int n = 1; n <= DR + DT; n++) {
    zp = tmul(&zp, &z);
    r = tadd(r, tscale(zp, coef[n]));
  }
  return r;
}
static tay trecip(const tay *a) {
  double a0 = a->c[0][0], coef[DR + DT + 1], p = 1.0 / a0;
  for (int n = 0; n <= DR + DT; n++) {
    coef[n] = p;
    p *= -1.0 / a0;
  }
  return tcompose(a, coef);
}
static tay tsqrt(const tay *a) {
  double a0 = a->c[0][0], coef[DR + DT + 1];
  double s = sqrt(a0), bin = 1.0, ia = 1.0 / a0, ipow = 1.0;
  for (int n = 0; n <= DR + DT; n++) {
    coef[n] = s * bin * ipow;
    bin *= (0.5 - n) / (n + 1);
    ipow *= ia;
  }
  return tcompose(a, coef);
}
/* Fermi function sigma(u) = 1/(exp(u)+1) of a Taylor argument.  Its
 * derivatives are polynomials in p = sigma, q = 1 - sigma with
 * dp/du = -p q, dq/du = p q, so d/du (p^a q^b) = -a p^a q^(b+1) + b p^(a+1) q^b;
 * p and q are formed without cancellation and every term is bounded. */
static tay tfermi(const tay *u) {
  const int NO = DR + DT;
  double u0 = u->c[0][0], p, q;
  if (u0 > 0.0) {
    double e = exp(-u0);
    p = e / (1.0 + e);
    q = 1.0 / (1.0 + e);
  } else {
    double e = exp(u0);
    p = 1.0 / (1.0 + e);
    q = e / (1.0 + e);
  }
  /* poly[a][b]: coefficient of p^a q^b in the n-th derivative */
  double poly[NO + 2][NO + 2], nxt[NO + 2][NO + 2], coef[NO + 1], fact = 1.0;
  memset(poly, 0, sizeof poly);
  poly[1][0] = 1.0;
  double pp[NO + 2], qq[NO + 2];
  pp[0] = qq[0] = 1.0;
  for (int a = 1; a <= NO + 1; a++) {
    pp[a] = pp[a - 1] * p;
    qq[a] = qq[a - 1] * q;
  }
  for (int n = 0; n <= NO; n++) {
    double v = 0.0;
    for (int a = 0; a <= NO + 1; a++)
      for (int b = 0; b <= NO + 1; b++)
        if (poly[a][b] != 0.0) v += poly[a][b] * pp[a] * qq[b];
    if (n) fact *= n;
    coef[n] = v / fact;
    memset(nxt, 0, sizeof nxt);
    for (int a = 0; a <= NO; a++)
      for (int b = 0; b <= NO; b++) {
        if (poly[a][b] == 0.0) continue;
        nxt[a][b + 1] += -a * poly[a][b];
        nxt[a + 1][b] += b * poly[a][b];
      }
    memcpy(poly, nxt, sizeof poly);
  }
  return tcompose(u, coef);
}

static tay tdiv(const tay *a, const tay *b) {
  tay r = trecip(b);
  return tmul(a, &r);
}

/* ------------------------------------------------------------ quadrature */
/* Composite rule in the kinetic energy x (in kT).  The Fermi function has
 * poles at x = eta +- i pi (2k+1), so the panels around c = max(eta, 0) are
 * graded (widths 4..16) with NPAN points each; below c - 40 the integrand is
 * smooth and is integrated in u = sqrt(x) (removes the sqrt(x) endpoint
 * singularity) over NU_PAN equal u-panels; above c + 40 a Gauss-Laguerre
 * rule takes the exponential tail. */
#define NPAN 16
#define NU_PAN 4
#define NUPTS 24
#define NLAG 32
#define MAXNODES (NU_PAN * NUPTS + 10 * NPAN + NLAG)
static double glp_x[NPAN], glp_w[NPAN], glu_x[NUPTS], glu_w[NUPTS], lag_x[NLAG], lag_w[NLAG];

static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; i++) {
    double z = cos(PI_ * (i + 0.75) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 0; j < n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j + 1.0) * z * p2 - j * p3) / (j + 1);
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > 1e-15);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

static void gauss_laguerre(int n, double *x, double *w) {
  double z = 0.0;
  for (int i = 0; i < n; i++) {
    if (i == 0) z = 3.0 / (1.0 + 2.4 * n);
    else if (i == 1) z += 15.0 / (1.0 + 2.5 * n);
    else z += ((1.0 + 2.55 * (i - 1)) / (1.9 * (i - 1))) * (z - x[i - 2]);
    double p1, p2, pp, z1;
    for (int it = 0; it < 200; it++) {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 0; j < n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j + 1.0 - z) * p2 - j * p3) / (j + 1);
      }
      pp = n * (p1 - p2) / z;
      z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) <= 1e-14 * fabs(z)) break;
    }
    x[i] = z;
    /* weight of int e^{-x} g(x) dx; the e^{-x} factor is divided out later */
    w[i] = -1.0 / (pp * n * p2);
  }
}

/* Nodes and weights of int_0^inf g(x) dx for a Fermi function centred at
 * eta.  Returns the node count (<= MAXNODES). */
static int fermi_nodes(double eta, double *xs, double *ws) {
  static const double off[11] = {-40.0, -24.0, -14.0, -8.0, -4.0, 0.0,
                                 4.0,   8.0,   14.0,  24.0, 40.0};
  const double c = eta > 0.0 ? eta : 0.0;
  int m = 0;
  double lo = 0.0; /* left end of the next panel */
  int first = 1;   /* the panel touching x = 0 is done in u = sqrt(x) */
  if (c + off[0] > 0.0) {
    const double ua = sqrt(c + off[0]);
    for (int p = 0; p < NU_PAN; p++) {
      const double u0 = ua * p / NU_PAN, h = 0.5 * ua / NU_PAN;
      for (int q = 0; q < NUPTS; q++) {
        const double u = u0 + h * (glu_x[q] + 1.0);
        xs[m] = u * u;
        ws[m] = h * glu_w[q] * 2.0 * u;
        m++;
      }
    }
    lo = c + off[0];
    first = 0;
  }
  for (int k = 1; k < 11; k++) {
    const double hi = c + off[k];
    if (hi <= lo) continue;
    if (first || lo < hi - lo) {
      /* near x = 0 the sqrt(x) branch point is removed by u = sqrt(x) */
      const double ul = sqrt(lo), h = 0.5 * (sqrt(hi) - ul);
      for (int q = 0; q < NPAN; q++) {
        const double u = ul + h * (glp_x[q] + 1.0);
        xs[m] = u * u;
        ws[m] = h * glp_w[q] * 2.0 * u;
        m++;
      }
      first = 0;
    } else {
      const double h = 0.5 * (hi - lo);
      for (int q = 0; q < NPAN; q++) {
        xs[m] = lo + h * (glp_x[q] + 1.0);
        ws[m] = h * glp_w[q];
        m++;
      }
    }
    lo = hi;
  }
  for (int q = 0; q < NLAG; q++) {
    xs[m] = lo + lag_x[q];
    ws[m] = lag_w[q] * exp(lag_x[q]);
    m++;
  }
  return m;
}

/* The three integrals F_{1/2}, F_{3/2}, F_{5/2} of one species with Taylor
 * eta and beta: out[0..2]. */
static void fermi_dirac(const tay *eta, const tay *beta, tay out[3]) {
  const double e0 = eta->c[0][0];
  for (int k = 0; k < 3; k++) out[k] = tconst(0.0);
  if (e0 < -700.0) return; /* species absent (e.g. positrons in cold matter) */
  double xs[MAXNODES], ws[MAXNODES];
  const int m = fermi_nodes(e0, xs, ws);
  tay halfb = tscale(*beta, 0.5);
  for (int q = 0; q < m; q++) {
    const double x = xs[q], w = ws[q];
    /* integrand without x^k: sqrt(1 + beta x / 2) / (exp(x - eta) + 1) */
    tay s = tscale(halfb, x);
    s.c[0][0] += 1.0;
    s = tsqrt(&s);
    tay arg = tscale(*eta, -1.0);
    arg.c[0][0] += x;
    if (arg.c[0][0] > 700.0) continue; /* exp(-700) relative: negligible */
    tay fd = tfermi(&arg);
    tay g = tmul(&s, &fd);
    double xk = sqrt(x);
    for (int k = 0; k < 3; k++) {
      out[k] = tadd(out[k], tscale(g, w * xk));
      xk *= x;
    }
  }
}

/* Scalar version for the eta solve: F_{1/2}, F_{3/2} and their eta
 * derivatives for one species. */
static void fermi_dirac_scalar(double eta, double beta, double F[2], double dF[2]) {
  F[0] = F[1] = dF[0] = dF[1] = 0.0;
  if (eta < -700.0) return;
  double xs[MAXNODES], ws[MAXNODES];
  const int m = fermi_nodes(eta, xs, ws);
  for (int q = 0; q < m; q++) {
    const double x = xs[q], w = ws[q];
    if (x - eta > 700.0) continue;
    double ex = exp(x - eta);
    double fd = 1.0 / (ex + 1.0);
    double dfd = ex * fd * fd; /* d fd / d eta */
    double g = sqrt(1.0 + 0.5 * beta * x) * w * sqrt(x);
    F[0] += g * fd;
    dF[0] += g * dfd;
    F[1] += g * x * fd;
    dF[1] += g * x * dfd;
  }
}

typedef struct {
  double kn, mc2;
} consts;

/* net number density and d/deta at scalar (eta, beta) */
static void nnet_scalar(double eta, double beta, const consts *K, double *n, double *dn) {
  double Fe[2], dFe[2], Fp[2], dFp[2];
  fermi_dirac_scalar(eta, beta, Fe, dFe);
  fermi_dirac_scalar(-eta - 2.0 / beta, beta, Fp, dFp);
  double b32 = beta * sqrt(beta);
  *n = K->kn * b32 * ((Fe[0] + beta * Fe[1]) - (Fp[0] + beta * Fp[1]));
  /* d eta_p / d eta = -1 */
  *dn = K->kn * b32 * ((dFe[0] + beta * dFe[1]) + (dFp[0] + beta * dFp[1]));
}

/* net electron number density and total pair pressure */
static void ndens_pres(const tay *eta, const tay *beta, const consts *K, tay *nnet,
                       tay *ptot) {
  tay Fe[3], Fp[3];
  fermi_dirac(eta, beta, Fe);
  tay etap = tscale(*eta, -1.0);
  tay twob = trecip(beta);
  twob = tscale(twob, 2.0);
  etap = tsub(etap, twob);
  fermi_dirac(&etap, beta, Fp);
  tay b12 = tsqrt(beta); /* beta^{1/2} */
  tay b32 = tmul(beta, &b12), b52 = tmul(&b32, beta);
  tay t, u;
  /* n = Kn beta^{3/2} (F12 + beta F32) */
  t = tmul(beta, &Fe[1]);
  t = tadd(Fe[0], t);
  u = tmul(beta, &Fp[1]);
  u = tadd(Fp[0], u);
  t = tsub(t, u);
  *nnet = tscale(tmul(&b32, &t), K->kn);
  /* P = 2/3 Kn mc2 beta^{5/2} (F32 + beta/2 F52) */
  tay hb = tscale(*beta, 0.5);
  t = tmul(&hb, &Fe[2]);
  t = tadd(Fe[1], t);
  u = tmul(&hb, &Fp[2]);
  u = tadd(Fp[1], u);
  t = tadd(t, u);
  *ptot = tscale(tmul(&b52, &t), (2.0 / 3.0) * K->kn * K->mc2);
}

/* one table point: fills f9 (reference order f, fd, ft, fdd, ftt, fdt, fddt,
 * fdtt, fddtt) and c4 (c, cd, ct, cdt) with c = dP/d(rhoYe), and eta, n. */
static int point(double rhos, double T, const consts *K, double *f9, double *c4,
                 double *eta4, double *xn4) {
  tay beta = tconst(K_B * T / K->mc2);
  beta.c[0][1] = K_B / K->mc2;
  tay target = tconst(rhos * N_AV);
  target.c[1][0] = N_AV;
  /* scalar solve for eta0: safeguarded Newton (bisection fallback) */
  double lo = -300.0, hi = 1.0e12, e0;
  const double b0 = beta.c[0][0];
  {
    double nq = K->kn * pow(b0, 1.5);
    double ratio = target.c[0][0] / nq;
    e0 = ratio < 1.0 ? log(ratio / 0.886) : pow(1.5 * ratio, 2.0 / 3.0);
    if (e0 < lo) e0 = lo + 1.0;
    for (int it = 0; it < 400; it++) {
      double n, dn;
      nnet_scalar(e0, b0, K, &n, &dn);
      double r = n - target.c[0][0];
      if (r > 0.0) hi = e0; else lo = e0;
      double en = e0 - r / dn;
      if (!(en > lo && en < hi) || !isfinite(en)) en = (hi > 1e6 && lo > 0) ? sqrt(lo * hi) : 0.5 * (lo + hi);
      if (fabs(en - e0) <= 4e-16 * fabs(en) + 1e-300) {
        e0 = en;
        break;
      }
      e0 = en;
    }
  }
  /* Taylor Newton: eta <- eta - (n(eta) - target) / (dn/deta at the point) */
  tay eta = tconst(e0), nn, pp;
  double dnde, ndum;
  nnet_scalar(e0, b0, K, &ndum, &dnde);
  for (int it = 0; it < DR + DT + 2; it++) {
    ndens_pres(&eta, &beta, K, &nn, &pp);
    tay r = tsub(nn, target);
    eta = tsub(eta, tscale(r, 1.0 / dnde));
  }
  ndens_pres(&eta, &beta, K, &nn, &pp);
  if (getenv("HELMGEN_TRACE")) {
    const tay *ts[3] = {&eta, &nn, &pp};
    const char *nm[3] = {"eta", "n", "P"};
    for (int t = 0; t < 3; t++) {
      fprintf(stderr, "%s:", nm[t]);
      for (int a = 0; a <= DR; a++)
        for (int b = 0; b <= DT; b++) fprintf(stderr, " [%d%d]=%g", a, b, ts[t]->c[a][b]);
      fprintf(stderr, "\n");
    }
  }
  /* f = eta k T N_A - P / rhoYe */
  tay kT = tscale(beta, K->mc2);
  tay rs = tconst(rhos);
  rs.c[1][0] = 1.0;
  tay t1 = tmul(&eta, &kT);
  t1 = tscale(t1, N_AV);
  tay t2 = tdiv(&pp, &rs);
  tay f = tsub(t1, t2);
  /* derivatives: d^{a+b} f / drho^a dT^b = a! b! c[a][b] */
  const double fa[4] = {1.0, 1.0, 2.0, 6.0};
#define DER(a, b) (fa[a] * fa[b] * f.c[a][b])
  f9[0] = DER(0, 0);
  f9[1] = DER(1, 0);
  f9[2] = DER(0, 1);
  f9[3] = DER(2, 0);
  f9[4] = DER(0, 2);
  f9[5] = DER(1, 1);
  f9[6] = DER(2, 1);
  f9[7] = DER(1, 2);
  f9[8] = DER(2, 2);
  /* P = rhoYe^2 df/drhoYe; c = dP/drhoYe and its derivatives */
  tay fr;
  memset(&fr, 0, sizeof fr);
  for (int a = 0; a < DR; a++)
    for (int b = 0; b <= DT; b++) fr.c[a][b] = (a + 1) * f.c[a + 1][b];
  tay rs2 = tmul(&rs, &rs);
  tay P = tmul(&rs2, &fr);
  tay c;
  memset(&c, 0, sizeof c);
  for (int a = 0; a < DR; a++)
    for (int b = 0; b <= DT; b++) c.c[a][b] = (a + 1) * P.c[a + 1][b];
  c4[0] = c.c[0][0];
  c4[1] = c.c[1][0];
  c4[2] = c.c[0][1];
  c4[3] = c.c[1][1];
  eta4[0] = eta.c[0][0];
  eta4[1] = eta.c[1][0];
  eta4[2] = eta.c[0][1];
  eta4[3] = eta.c[1][1];
  xn4[0] = nn.c[0][0];
  xn4[1] = nn.c[1][0];
  xn4[2] = nn.c[0][1];
  xn4[3] = nn.c[1][1];
  for (int v = 0; v < 9; v++)
    if (!isfinite(f9[v])) return -1;
  for (int v = 0; v < 4; v++)
    if (!isfinite(c4[v])) return -1;
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 8) {
    fprintf(stderr, "usage: %s nrho nT ldlo ldhi ltlo lthi out.bin [out.dat]\n", argv[0]);
    return 2;
  }
  const int nr = atoi(argv[1]), nt = atoi(argv[2]);
  const double ldlo = atof(argv[3]), ldhi = atof(argv[4]);
  const double ltlo = atof(argv[5]), lthi = atof(argv[6]);
  gauss_legendre(NPAN, glp_x, glp_w);
  gauss_legendre(NUPTS, glu_x, glu_w);
  gauss_laguerre(NLAG, lag_x, lag_w);
  consts K;
  K.mc2 = M_ELEC * C_LIGHT * C_LIGHT;
  K.kn = 8.0 * PI_ * sqrt(2.0) * pow(M_ELEC * C_LIGHT / H_PL, 3.0);
  const long np = (long)nr * nt;
  double *f9 = (double *)malloc(sizeof(double) * 9 * np);
  double *c4 = (double *)malloc(sizeof(double) * 4 * np);
  double *e4 = (double *)malloc(sizeof(double) * 4 * np);
  double *x4 = (double *)malloc(sizeof(double) * 4 * np);
  /* the reference's node placement: helm_T(j) = 10**(ltlo + (j-1) dT),
   * helm_rho(i) = 10**(ldlo + (i-1) drho)  (source/source.F90:23094-23104) */
  const double dT = (lthi - ltlo) / (double)(nt - 1);
  const double dR = (ldhi - ldlo) / (double)(nr - 1);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
  for (long p = 0; p < np; p++) {
    int j = (int)(p / nr), i = (int)(p % nr);
    double T = pow(10.0, ltlo + (double)j * dT);
    double rho = pow(10.0, ldlo + (double)i * dR);
    if (point(rho, T, &K, f9 + 9 * p, c4 + 4 * p, e4 + 4 * p, x4 + 4 * p)) {
      bad++;
      if (getenv("HELMGEN_VERBOSE"))
      {
        fprintf(stderr, "bad point logrho=%.3f logT=%.3f eta=%g:", log10(rho), log10(T), e4[4 * p]);
        for (int v = 0; v < 9; v++) fprintf(stderr, " %g", f9[9 * p + v]);
        fprintf(stderr, " |");
        for (int v = 0; v < 4; v++) fprintf(stderr, " %g", c4[4 * p + v]);
        fprintf(stderr, "\n");
      }
    }
  }
  if (bad) fprintf(stderr, "helmgen: %d non-finite points\n", bad);
  FILE *fb = fopen(argv[7], "wb");
  if (!fb) return 1;
  const char magic[8] = "PHLXHLM1";
  int dims[2] = {nr, nt};
  double box[4] = {ldlo, ldhi, ltlo, lthi};
  fwrite(magic, 1, 8, fb);
  fwrite(dims, sizeof(int), 2, fb);
  fwrite(box, sizeof(double), 4, fb);
  fwrite(f9, sizeof(double), 9 * np, fb);
  fwrite(c4, sizeof(double), 4 * np, fb);
  fclose(fb);
  if (argc > 8) {
    FILE *fa = fopen(argv[8], "w");
    if (!fa) return 1;
    for (long p = 0; p < np; p++) {
      for (int v = 0; v < 9; v++) fprintf(fa, "%.17e%s", f9[9 * p + v], v == 8 ? "\n" : " ");
    }
    for (long p = 0; p < np; p++)
      fprintf(fa, "%.17e %.17e %.17e %.17e\n", c4[4 * p], c4[4 * p + 1], c4[4 * p + 2], c4[4 * p + 3]);
    for (long p = 0; p < np; p++)
      fprintf(fa, "%.17e %.17e %.17e %.17e\n", e4[4 * p], e4[4 * p + 1], e4[4 * p + 2], e4[4 * p + 3]);
    for (long p = 0; p < np; p++)
      fprintf(fa, "%.17e %.17e %.17e %.17e\n", x4[4 * p], x4[4 * p + 1], x4[4 * p + 2], x4[4 * p + 3]);
    fclose(fa);
  }
  free(f9); free(c4); free(e4); free(x4);
  return bad ? 1 : 0;
}
