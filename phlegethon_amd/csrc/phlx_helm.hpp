// phlx_helm.hpp -- device Helmholtz EoS: biquintic Hermite interpolation of
// the electron-positron free energy f(rho*Ye, T), bicubic dP/d(rho*Ye), ions
// and radiation, and the Newton inversions T(rho, e) and T(rho, P).
//
// Reference: helm_rhoT_given_full (source/source.F90:23686-24232),
// helm_rhoT_given1 (:24234-24569), helm_rhoT_given2 (:24571-24941),
// helm_rhoP_given (:24943-24980), helm_rhoe_given (:24982-25019).
// Coulomb corrections are off in the reference build and are not carried.
//
// Table layout in HBM: one 128-byte record per node, [jT][irho][16]:
//   0 f, 1 df/dT, 2 d2f/dT2, 3 df/drho, 4 d2f/drho2, 5 d2f/drhodT,
//   6 d3f/drho2dT, 7 d3f/drhodT2, 8 d4f/drho2dT2,
//   9 c, 10 dc/dT, 11 dc/drho, 12 d2c/drhodT   (c = dP/d(rho*Ye)), 13-15 pad.
// The four corner records of a cell are two adjacent pairs, so an evaluation
// touches four 128-byte lines; the whole 541x271 table (18.8 MB) stays in the
// last-level cache once the working set of a sweep has touched it.
// Arithmetic is written term by term in the reference's order (no
// contraction) so that results are reproducible bit for bit on any host.
#pragma once

#include <hip/hip_runtime.h>

namespace phlx {

constexpr double kHelmRgas = 8.31446261815324e7;  // source/source.F90:605-616
constexpr double kHelmCrad = 7.565767381646406e-15;
constexpr double kHelmClight = 2.99792458e10;
constexpr double kHelmClight2 = 8.987551787368177e20;
constexpr int kHelmRec = 16;
constexpr int kHelmMaxIter = 100;   // source/source.F90:24955, :24994
constexpr double kHelmTol = 1e-11;  // source/source.F90:754

struct HelmDev {
  const double *tab;  // [nt][nr][16]
  const double *rho;  // node values, nr
  const double *T;    // node values, nt
  int nr, nt;
  double lrho1, lT1;  // log10 of the first nodes
  double dr, dt;      // log10 spacings
  double tl, th;      // Tl_eos, Th_eos (source/source.F90:750-751)
  // PIG_EOS table (pig_rhoT_given_*, source/source.F90:25457-26731): the free
  // energy of the whole gas at rho, i.e. Ye = 1 and no ion terms; the
  // evaluations below then take Ye = 1 and 1/abar = 0, which leaves the
  // reference's PIG arithmetic unchanged bit for bit (x * 1 = x, x + 0 = x)
  int pig;
};

struct HelmCell {
  int ih, jh;
  double x, y, drho, dT;
};

// Marks v as computed at this point.  Without it the compiler sinks a whole
// evaluation into the conditional stores that follow (one copy per branch),
// which holds every table value of the evaluation live at once.
__device__ __forceinline__ void helm_pin(double &v) { asm volatile("" : "+v"(v)); }

// a / b; FAST (the mixed-precision EoS step only): the hardware reciprocal
// refined by two Newton steps (~1 ulp) instead of the IEEE-rounded sequence.
template <bool FAST>
__device__ __forceinline__ double hdiv(double a, double b) {
  if (!FAST) return a / b;
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  return a * r;
}

// Basis values of one coordinate at the cell's low corner (a[0..2]: order
// 0, 1, 2 of the Hermite family) and high corner (am[0..2]).
struct HelmBasis {
  double a[3], am[3];
};

// cell index int((log10(rhoYe) - log10(rho_1)) / dlog) (source/source.F90
// :23773-23784), kept inside the table.  FAST (the mixed-precision EoS step
// only): the index from the FP32 base-2 logarithm (v_log_f32) instead of two
// FP64 log10; within ~1e-5 of a node it may pick the neighbouring cell, whose
// quintic Hermite interpolant matches value, slope and curvature there, and
// the offsets x, y are still formed in FP64 from the table's nodes.
template <bool FAST = false>
__device__ __forceinline__ HelmCell helm_locate(const HelmDev &h, double rhos, double T) {
  HelmCell c;
  if (FAST) {
    constexpr float kLog10of2 = 0.30102999566398120f;
    const float lr = __builtin_amdgcn_logf((float)rhos) * kLog10of2;
    const float lt = __builtin_amdgcn_logf((float)T) * kLog10of2;
    c.ih = (int)((lr - (float)h.lrho1) * __builtin_amdgcn_rcpf((float)h.dr));
    c.jh = (int)((lt - (float)h.lT1) * __builtin_amdgcn_rcpf((float)h.dt));
  } else {
    c.ih = (int)((log10(rhos) - h.lrho1) / h.dr);
    c.jh = (int)((log10(T) - h.lT1) / h.dt);
  }
  c.ih = c.ih < 0 ? 0 : (c.ih > h.nr - 2 ? h.nr - 2 : c.ih);
  c.jh = c.jh < 0 ? 0 : (c.jh > h.nt - 2 ? h.nt - 2 : c.jh);
  double tmp = h.rho[c.ih];
  c.drho = h.rho[c.ih + 1] - tmp;
  c.x = hdiv<FAST>(rhos - tmp, c.drho);
  tmp = h.T[c.jh];
  c.dT = h.T[c.jh + 1] - tmp;
  c.y = hdiv<FAST>(T - tmp, c.dT);
  return c;
}

// Quintic Hermite basis of one coordinate and its first / second derivative
// (source/source.F90:23832-23862, :23900-23906, :24030-24036).
__device__ __forceinline__ void helm_basis(double x, double d, HelmBasis &v) {
  const double d2 = d * d;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x;
  const double o = 1.0 - x;
  const double o2 = o * o, o3 = o2 * o, o4 = o3 * o, o5 = o4 * o;
  v.a[0] = -6.0 * x5 + 15.0 * x4 - 10.0 * x3 + 1.0;
  v.a[1] = (-3.0 * x5 + 8.0 * x4 - 6.0 * x3 + x) * d;
  v.a[2] = (0.5 * (-x5 + 3.0 * x4 - 3.0 * x3 + x2)) * d2;
  v.am[0] = -6.0 * o5 + 15.0 * o4 - 10.0 * o3 + 1.0;
  v.am[1] = (3.0 * o5 - 8.0 * o4 + 6.0 * o3 - o) * d;
  v.am[2] = (0.5 * (-o5 + 3.0 * o4 - 3.0 * o3 + o2)) * d2;
}

template <bool FAST = false>
__device__ __forceinline__ void helm_dbasis(double x, double d, HelmBasis &v) {
  const double id = hdiv<FAST>(1.0, d);
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x;
  const double o = 1.0 - x;
  const double o2 = o * o, o3 = o2 * o, o4 = o3 * o;
  v.a[0] = (-30.0 * x4 + 60.0 * x3 - 30.0 * x2) * id;
  v.a[1] = (-15.0 * x4 + 32.0 * x3 - 18.0 * x2 + 1.0);
  v.a[2] = (0.5 * (-5.0 * x4 + 12.0 * x3 - 9.0 * x2 + 2.0 * x)) * d;
  v.am[0] = -(-30.0 * o4 + 60.0 * o3 - 30.0 * o2) * id;
  v.am[1] = (-15.0 * o4 + 32.0 * o3 - 18.0 * o2 + 1.0);
  v.am[2] = -(0.5 * (-5.0 * o4 + 12.0 * o3 - 9.0 * o2 + 2.0 * o)) * d;
}

template <bool FAST = false>
__device__ __forceinline__ void helm_ddbasis(double x, double d, HelmBasis &v) {
  const double id = hdiv<FAST>(1.0, d), id2 = id * id;
  const double x2 = x * x, x3 = x2 * x;
  const double o = 1.0 - x;
  const double o2 = o * o, o3 = o2 * o;
  v.a[0] = (-120.0 * x3 + 180.0 * x2 - 60.0 * x) * id2;
  v.a[1] = (-60.0 * x3 + 96.0 * x2 - 36.0 * x) * id;
  v.a[2] = (0.5 * (-20.0 * x3 + 36.0 * x2 - 18.0 * x + 2.0));
  v.am[0] = (-120.0 * o3 + 180.0 * o2 - 60.0 * o) * id2;
  v.am[1] = -(-60.0 * o3 + 96.0 * o2 - 36.0 * o) * id;
  v.am[2] = (0.5 * (-20.0 * o3 + 36.0 * o2 - 18.0 * o + 2.0));
}

// NS biquintic sums over the cell in ONE pass over the table: sum k pairs
// the rho basis R[k] with the T basis T[k].  Each sum keeps the reference's
// term order: entry v pairs the rho basis of order (0,0,0,1,2,1,2,1,2)[v]
// with the T basis of order (0,1,2,0,0,1,1,2,2)[v]; per entry the corners
// (i,j) (i+1,j) (i,j+1) (i+1,j+1) are added in that order.  Reading each
// table value once for all sums keeps ~4 values in flight per lane instead
// of the whole 36-value block per sum.
// FMA (the mixed-precision EoS step only): the basis products are formed
// once per entry and the four corner terms accumulated with fused
// multiply-adds (8 instead of 12 operations per entry and sum); the FP64
// path keeps the reference's unfused order.
template <int NS, bool FMA = false>
__device__ __forceinline__ void helm_sums36(const HelmDev &h, const HelmCell &c,
                                            const HelmBasis *const (&R)[NS],
                                            const HelmBasis *const (&T)[NS], double (&s)[NS]) {
  const double *n00 = h.tab + (size_t)kHelmRec * ((size_t)c.jh * h.nr + c.ih);
  const double *n10 = n00 + kHelmRec;
  const double *n01 = n00 + (size_t)kHelmRec * h.nr;
  const double *n11 = n01 + kHelmRec;
  constexpr int RX[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2};
  constexpr int TY[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
#pragma unroll
  for (int v = 0; v < 9; v++) {
    const double q0 = n00[v], q1 = n10[v], q2 = n01[v], q3 = n11[v];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const double a = R[k]->a[RX[v]], am = R[k]->am[RX[v]];
      const double b = T[k]->a[TY[v]], bm = T[k]->am[TY[v]];
      if (FMA) {
        const double ab = a * b, amb = am * b, abm = a * bm, ambm = am * bm;
        s[k] = (v == 0) ? q0 * ab : fma(q0, ab, s[k]);
        s[k] = fma(q1, amb, s[k]);
        s[k] = fma(q2, abm, s[k]);
        s[k] = fma(q3, ambm, s[k]);
      } else {
        const double t0 = q0 * a * b, t1 = q1 * am * b;
        const double t2 = q2 * a * bm, t3 = q3 * am * bm;
        s[k] = (v == 0) ? t0 : s[k] + t0;
        s[k] = s[k] + t1;
        s[k] = s[k] + t2;
        s[k] = s[k] + t3;
      }
    }
  }
}

// Mixed-precision EoS only: the same biquintic sums contracted separably.
// The cell's 36 table values times the rho basis R_k are first summed over
// the rho order and corner, G_k[ty][yc] = sum_{rx, xc} q(rx, ty; xc, yc)
// R_k(rx, xc) (9 entries x 4 corners), then every sum pairing R_k with a T
// basis is sum_{ty, yc} G_k[ty][yc] T(ty, yc) (6 terms): 36 + 6 fused
// multiply-adds per sum where helm_sums36 takes 72 operations.  The terms
// are added in another order than the reference's (round-off level).
template <int NR>
__device__ __forceinline__ void helm_contract_rho(const HelmDev &h, const HelmCell &c,
                                                  const HelmBasis *const (&R)[NR],
                                                  double (&G)[NR][6]) {
  const double *n00 = h.tab + (size_t)kHelmRec * ((size_t)c.jh * h.nr + c.ih);
  const double *n10 = n00 + kHelmRec;
  const double *n01 = n00 + (size_t)kHelmRec * h.nr;
  const double *n11 = n01 + kHelmRec;
  constexpr int RX[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2};
  constexpr int TY[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
#pragma unroll
  for (int v = 0; v < 9; v++) {
    const double q0 = n00[v], q1 = n10[v], q2 = n01[v], q3 = n11[v];
    const int lo = 2 * TY[v], hi = lo + 1;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const double a = R[k]->a[RX[v]], am = R[k]->am[RX[v]];
      // entries 0..2 (rho order 0) are the first of their T order
      G[k][lo] = (v < 3) ? q0 * a : fma(q0, a, G[k][lo]);
      G[k][lo] = fma(q1, am, G[k][lo]);
      G[k][hi] = (v < 3) ? q2 * a : fma(q2, a, G[k][hi]);
      G[k][hi] = fma(q3, am, G[k][hi]);
    }
  }
}
__device__ __forceinline__ double helm_contract_T(const double (&G)[6], const HelmBasis &T) {
  double s = G[0] * T.a[0];
  s = fma(G[1], T.am[0], s);
  s = fma(G[2], T.a[1], s);
  s = fma(G[3], T.am[1], s);
  s = fma(G[4], T.a[2], s);
  s = fma(G[5], T.am[2], s);
  return s;
}

// Bicubic Hermite of the dP/d(rho*Ye) block (source/source.F90:24078-24125).
template <bool FMA = false>
__device__ __forceinline__ double helm_sum16(const HelmDev &h, const HelmCell &c) {
  const double x = c.x, y = c.y, o = 1.0 - x, oy = 1.0 - y;
  const double x2 = x * x, x3 = x2 * x, o2 = o * o, o3 = o2 * o;
  const double y2 = y * y, y3 = y2 * y, oy2 = oy * oy, oy3 = oy2 * oy;
  const double p0r = 2.0 * x3 - 3.0 * x2 + 1.0;
  const double p1r = (x3 - 2.0 * x2 + x) * c.drho;
  const double p0mr = 2.0 * o3 - 3.0 * o2 + 1.0;
  const double p1mr = -(o3 - 2.0 * o2 + o) * c.drho;
  const double p0t = 2.0 * y3 - 3.0 * y2 + 1.0;
  const double p1t = (y3 - 2.0 * y2 + y) * c.dT;
  const double p0mt = 2.0 * oy3 - 3.0 * oy2 + 1.0;
  const double p1mt = -(oy3 - 2.0 * oy2 + oy) * c.dT;
  const double *n00 = h.tab + (size_t)kHelmRec * ((size_t)c.jh * h.nr + c.ih) + 9;
  const double *n10 = n00 + kHelmRec;
  const double *n01 = n00 + (size_t)kHelmRec * h.nr;
  const double *n11 = n01 + kHelmRec;
  const double rb[4] = {p0r, p0r, p1r, p1r}, rbm[4] = {p0mr, p0mr, p1mr, p1mr};
  const double tb[4] = {p0t, p1t, p0t, p1t}, tbm[4] = {p0mt, p1mt, p0mt, p1mt};
  double s = 0.0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    if (FMA) {
      s = (v == 0) ? n00[v] * (rb[v] * tb[v]) : fma(n00[v], rb[v] * tb[v], s);
      s = fma(n10[v], rbm[v] * tb[v], s);
      s = fma(n01[v], rb[v] * tbm[v], s);
      s = fma(n11[v], rbm[v] * tbm[v], s);
    } else {
      const double t0 = n00[v] * rb[v] * tb[v], t1 = n10[v] * rbm[v] * tb[v];
      const double t2 = n01[v] * rb[v] * tbm[v], t3 = n11[v] * rbm[v] * tbm[v];
      s = (v == 0) ? t0 : s + t0;
      s = s + t1;
      s = s + t2;
      s = s + t3;
    }
  }
  return s;
}

// helm_rhoT_given_full: P, E (specific), sound speed (relativistic form), c_v;
// dPdT (when non-null) receives dP/dT.  FMA: fused table sums (mixed path).
template <bool FMA = false>
__device__ __forceinline__ void helm_full(const HelmDev &h, double rho, double T, double abar,
                                          double zbar, double &P, double &E, double &sound,
                                          double &cv, double *dPdT_out = nullptr) {
  const double Ye = h.pig ? 1.0 : hdiv<FMA>(zbar, abar), rhos = rho * Ye;
  const HelmCell c = helm_locate<FMA>(h, rhos, T);
  HelmBasis r, dr, t, dt, ddt;
  helm_basis(c.x, c.drho, r);
  helm_dbasis<FMA>(c.x, c.drho, dr);
  helm_basis(c.y, c.dT, t);
  helm_dbasis<FMA>(c.y, c.dT, dt);
  helm_ddbasis<FMA>(c.y, c.dT, ddt);
  double sm[5];
  if constexpr (FMA) {
    const HelmBasis *const RR[2] = {&r, &dr};
    double G[2][6];
    helm_contract_rho<2>(h, c, RR, G);
    sm[0] = helm_contract_T(G[0], t);
    sm[1] = helm_contract_T(G[1], t);
    sm[2] = helm_contract_T(G[0], dt);
    sm[3] = helm_contract_T(G[1], dt);
    sm[4] = helm_contract_T(G[0], ddt);
  } else {
    const HelmBasis *const RR[5] = {&r, &dr, &r, &dr, &r};
    const HelmBasis *const TT[5] = {&t, &t, &dt, &dt, &ddt};
    helm_sums36<5, FMA>(h, c, RR, TT, sm);
  }
  const double f = sm[0], df_drho = sm[1], df_dT = sm[2], df_drhodT = sm[3], df_dT2 = sm[4];
  double dP_drho = helm_sum16<FMA>(h, c);
  dP_drho = Ye * dP_drho;
  double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double s = -Ye * df_dT;
  const double ds_dT = -Ye * df_dT2;
  double e = Ye * f + T * s;
  double dE_dT = T * ds_dT;
  tmp = h.pig ? 0.0 : hdiv<FMA>(1.0, abar);
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double pion = kHelmRgas * rho * T * tmp;
  const double dpiondt = kHelmRgas * rho * tmp;
  const double dpiondd = kHelmRgas * T * tmp;
  p = p + pion + kHelmCrad * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + dpiondt + (4.0 / 3.0) * kHelmCrad * T3;
  dP_drho = dP_drho + dpiondd;
  e = e + 1.5 * kHelmRgas * T * tmp + hdiv<FMA>(kHelmCrad * T4, rho);
  dE_dT = dE_dT + 1.5 * kHelmRgas * tmp + hdiv<FMA>(4.0 * kHelmCrad * T3, rho);
  const double zz = hdiv<FMA>(p, rho);
  const double chiT = hdiv<FMA>(T, p) * dP_dT;
  const double chirho = hdiv<FMA>(dP_drho, zz);
  const double x = hdiv<FMA>(zz * chiT, T * dE_dT);
  const double gam1 = chiT * x + chirho;
  const double z = 1.0 + hdiv<FMA>(e + kHelmClight2, zz);
  P = p;
  E = e;
  cv = dE_dT;
  sound = kHelmClight * sqrt(hdiv<FMA>(gam1, z));
  if (dPdT_out) *dPdT_out = dP_dT;
}

// ---------------------------------------------------------------------------
// Mixed-precision EoS: one cell of the table contracted over rho, reused by
// every Newton iterate of a solve.  Within a cell the interpolant is a
// polynomial in the T coordinate y (quintic for f, cubic for dP/drho), so
// once the 36 (16) table values are summed against the rho basis, f and its
// derivatives at any T are 6 (4) products with the T basis; the evaluation
// is the same interpolant in FP64, the terms added in another order than the
// reference's.  A Newton iterate that leaves the cell by more than 1 % of
// its width relocates; closer iterates keep the polynomial, whose difference
// from the neighbouring cell's is third order in the distance (the quintic
// Hermite interpolants share value, slope and curvature at the node).
struct HelmMx {
  int ih, jh;
  double x, drho, dT, Tj;
  double Ye, rhos, iabar;          // Ye, rho Ye, 1/abar (0: PIG)
  double Gr[6], Gdr[6], Hc[4];     // f against the rho basis and its slope; dP/drho block
};

__device__ __forceinline__ void helm_mx_contract(const HelmDev &h, double T, HelmMx &m) {
  const HelmCell c = helm_locate<true>(h, m.rhos, T);
  m.ih = c.ih;
  m.jh = c.jh;
  m.x = c.x;
  m.drho = c.drho;
  m.dT = c.dT;
  m.Tj = h.T[c.jh];
  HelmBasis r, dr;
  helm_basis(c.x, c.drho, r);
  helm_dbasis<true>(c.x, c.drho, dr);
  {
    const HelmBasis *const RR[2] = {&r, &dr};
    double G[2][6];
    helm_contract_rho<2>(h, c, RR, G);
#pragma unroll
    for (int q = 0; q < 6; q++) {
      m.Gr[q] = G[0][q];
      m.Gdr[q] = G[1][q];
    }
  }
  // bicubic dP/d(rho Ye) block: entries c, dc/dT, dc/drho, d2c/drhodT pair
  // the rho basis of order (0, 0, 1, 1) with the T basis of order (0, 1, 0, 1)
  const double x = c.x, o = 1.0 - x;
  const double x2 = x * x, x3 = x2 * x, o2 = o * o, o3 = o2 * o;
  const double rb[2] = {2.0 * x3 - 3.0 * x2 + 1.0, (x3 - 2.0 * x2 + x) * c.drho};
  const double rbm[2] = {2.0 * o3 - 3.0 * o2 + 1.0, -(o3 - 2.0 * o2 + o) * c.drho};
  const double *n00 = h.tab + (size_t)kHelmRec * ((size_t)c.jh * h.nr + c.ih) + 9;
  const double *n10 = n00 + kHelmRec;
  const double *n01 = n00 + (size_t)kHelmRec * h.nr;
  const double *n11 = n01 + kHelmRec;
  constexpr int RXc[4] = {0, 0, 1, 1}, TYc[4] = {0, 1, 0, 1};
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int lo = 2 * TYc[v], hi = lo + 1;
    const double a = rb[RXc[v]], am = rbm[RXc[v]];
    m.Hc[lo] = (v < 2) ? n00[v] * a : fma(n00[v], a, m.Hc[lo]);
    m.Hc[lo] = fma(n10[v], am, m.Hc[lo]);
    m.Hc[hi] = (v < 2) ? n01[v] * a : fma(n01[v], a, m.Hc[hi]);
    m.Hc[hi] = fma(n11[v], am, m.Hc[hi]);
  }
}

__device__ __forceinline__ void helm_mx_setup(const HelmDev &h, double rho, double T, double abar,
                                              double zbar, HelmMx &m) {
  m.Ye = h.pig ? 1.0 : hdiv<true>(zbar, abar);
  m.rhos = rho * m.Ye;
  m.iabar = h.pig ? 0.0 : hdiv<true>(1.0, abar);
  helm_mx_contract(h, T, m);
}

// E and c_v at T from a contracted cell (helm_mx_energy), and P, dP/dT and
// the sound speed at the same T (helm_mx_pressure, from the T basis the
// first kept in a HelmMxAt): the operations of helm_full / helm_given2
// after their sums (ions, radiation, relativistic sound speed).  A Newton
// iterate needs only the first; the solve's last iterate takes both.
struct HelmMxAt {
  double y;
  HelmBasis t, dt;
};

__device__ __forceinline__ void helm_mx_energy(const HelmDev &h, HelmMx &m, double rho, double T,
                                               HelmMxAt &a, double &E, double &cv) {
  double y = hdiv<true>(T - m.Tj, m.dT);
  if (y < -0.01 || y > 1.01) {
    helm_mx_contract(h, T, m);
    y = hdiv<true>(T - m.Tj, m.dT);
  }
  a.y = y;
  HelmBasis ddt;
  helm_basis(y, m.dT, a.t);
  helm_dbasis<true>(y, m.dT, a.dt);
  helm_ddbasis<true>(y, m.dT, ddt);
  const double f = helm_contract_T(m.Gr, a.t);
  const double df_dT = helm_contract_T(m.Gr, a.dt);
  const double df_dT2 = helm_contract_T(m.Gr, ddt);
  const double Ye = m.Ye, tmp = m.iabar;
  const double s = -Ye * df_dT;
  const double ds_dT = -Ye * df_dT2;
  double e = Ye * f + T * s;
  double dE_dT = T * ds_dT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  e = e + 1.5 * kHelmRgas * T * tmp + hdiv<true>(kHelmCrad * T4, rho);
  dE_dT = dE_dT + 1.5 * kHelmRgas * tmp + hdiv<true>(4.0 * kHelmCrad * T3, rho);
  E = e;
  cv = dE_dT;
}

__device__ __forceinline__ void helm_mx_pressure(const HelmMx &m, double rho, double T,
                                                 const HelmMxAt &a, double E, double cv,
                                                 double &P, double &dPdT, double &sound) {
  const double y = a.y, Ye = m.Ye, tmp = m.iabar;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double df_drho = helm_contract_T(m.Gdr, a.t);
  const double df_drhodT = helm_contract_T(m.Gdr, a.dt);
  const double oy = 1.0 - y, y2 = y * y, y3 = y2 * y, oy2 = oy * oy, oy3 = oy2 * oy;
  double c3 = m.Hc[0] * (2.0 * y3 - 3.0 * y2 + 1.0);
  c3 = fma(m.Hc[1], 2.0 * oy3 - 3.0 * oy2 + 1.0, c3);
  c3 = fma(m.Hc[2], (y3 - 2.0 * y2 + y) * m.dT, c3);
  c3 = fma(m.Hc[3], -(oy3 - 2.0 * oy2 + oy) * m.dT, c3);
  double dP_drho = Ye * c3;
  const double rr = m.rhos * m.rhos;
  double p = rr * df_drho;
  double dP_dT = rr * df_drhodT;
  const double pion = kHelmRgas * rho * T * tmp;
  const double dpiondt = kHelmRgas * rho * tmp;
  const double dpiondd = kHelmRgas * T * tmp;
  p = p + pion + kHelmCrad * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + dpiondt + (4.0 / 3.0) * kHelmCrad * T3;
  dP_drho = dP_drho + dpiondd;
  const double zz = hdiv<true>(p, rho);
  const double chiT = hdiv<true>(T, p) * dP_dT;
  const double chirho = hdiv<true>(dP_drho, zz);
  const double xx = hdiv<true>(zz * chiT, T * cv);
  const double gam1 = chiT * xx + chirho;
  const double z = 1.0 + hdiv<true>(E + kHelmClight2, zz);
  P = p;
  dPdT = dP_dT;
  sound = kHelmClight * sqrt(hdiv<true>(gam1, z));
}

// helm_rhoT_given1: P and (when ret) dP/dT.
__device__ __forceinline__ void helm_given1(const HelmDev &h, double rho, double T, double abar,
                                            double zbar, double &P, double &dPdT, bool ret) {
  const double Ye = h.pig ? 1.0 : zbar / abar, rhos = rho * Ye;
  const HelmCell c = helm_locate(h, rhos, T);
  HelmBasis dr, t, dt;
  helm_dbasis(c.x, c.drho, dr);
  helm_basis(c.y, c.dT, t);
  double df_drho, df_drhodT = 0.0;
  if (ret) {
    helm_dbasis(c.y, c.dT, dt);
    const HelmBasis *const RR[2] = {&dr, &dr};
    const HelmBasis *const TT[2] = {&t, &dt};
    double sm[2];
    helm_sums36<2>(h, c, RR, TT, sm);
    df_drho = sm[0];
    df_drhodT = sm[1];
  } else {
    const HelmBasis *const RR[1] = {&dr};
    const HelmBasis *const TT[1] = {&t};
    double sm[1];
    helm_sums36<1>(h, c, RR, TT, sm);
    df_drho = sm[0];
  }
  double tmp = rhos * rhos;
  double p = tmp * df_drho;
  double dP_dT = tmp * df_drhodT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  tmp = h.pig ? 0.0 : 1.0 / abar;
  const double pion = kHelmRgas * rho * T * tmp;
  const double dpiondt = kHelmRgas * rho * tmp;
  p = p + pion + kHelmCrad * T4 * (1.0 / 3.0);
  dP_dT = dP_dT + dpiondt + (4.0 / 3.0) * kHelmCrad * T3;
  P = p;
  dPdT = dP_dT;
}

// helm_rhoT_given2: E and (when ret) dE/dT.  FMA (the mixed-precision EoS
// only): the separable contraction, refined reciprocals, FP32 cell lookup.
template <bool FMA = false>
__device__ __forceinline__ void helm_given2(const HelmDev &h, double rho, double T, double abar,
                                            double zbar, double &E, double &dEdT, bool ret) {
  const double Ye = h.pig ? 1.0 : hdiv<FMA>(zbar, abar), rhos = rho * Ye;
  const HelmCell c = helm_locate<FMA>(h, rhos, T);
  HelmBasis r, t, dt, ddt;
  helm_basis(c.x, c.drho, r);
  helm_basis(c.y, c.dT, t);
  helm_dbasis<FMA>(c.y, c.dT, dt);
  double f, df_dT, df_dT2 = 0.0;
  if (FMA) {
    helm_ddbasis<FMA>(c.y, c.dT, ddt);
    const HelmBasis *const RR[1] = {&r};
    double G[1][6];
    helm_contract_rho<1>(h, c, RR, G);
    f = helm_contract_T(G[0], t);
    df_dT = helm_contract_T(G[0], dt);
    df_dT2 = helm_contract_T(G[0], ddt);
  } else if (ret) {
    helm_ddbasis(c.y, c.dT, ddt);
    const HelmBasis *const RR[3] = {&r, &r, &r};
    const HelmBasis *const TT[3] = {&t, &dt, &ddt};
    double sm[3];
    helm_sums36<3>(h, c, RR, TT, sm);
    f = sm[0];
    df_dT = sm[1];
    df_dT2 = sm[2];
  } else {
    const HelmBasis *const RR[2] = {&r, &r};
    const HelmBasis *const TT[2] = {&t, &dt};
    double sm[2];
    helm_sums36<2>(h, c, RR, TT, sm);
    f = sm[0];
    df_dT = sm[1];
  }
  const double s = -Ye * df_dT;
  const double ds_dT = -Ye * df_dT2;
  double e = Ye * f + T * s;
  double dE_dT = T * ds_dT;
  const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  const double tmp = h.pig ? 0.0 : hdiv<FMA>(1.0, abar);
  e = e + 1.5 * kHelmRgas * T * tmp + hdiv<FMA>(kHelmCrad * T4, rho);
  dE_dT = dE_dT + 1.5 * kHelmRgas * tmp + hdiv<FMA>(4.0 * kHelmCrad * T3, rho);
  E = e;
  dEdT = dE_dT;
}

// helm_rhoe_given: Newton for T(rho, e) from the warm start T (in/out).  The
// residual of the incoming T is tested after the update, so the returned T is
// one Newton step past the converged one; a T outside [Tl, Th] resets to the
// start value and ends the loop.  Returns the iterations taken (<= 100).
// cv_last (optional): dE/dT of the last Newton evaluation (the energy
// predictor's slope for the next solve of the zone).
__device__ __forceinline__ int helm_rhoe_given(const HelmDev &h, double rho, double E0,
                                               double abar, double zbar, double &T, double &P,
                                               double &sound, bool ret,
                                               double *cv_last = nullptr) {
  const double Told = T;
  int iter;
  sound = 0.0;
  for (iter = 1; iter <= kHelmMaxIter; iter++) {
    double e, dEdT;
    helm_given2(h, rho, T, abar, zbar, e, dEdT, true);
    if (cv_last) *cv_last = dEdT;
    const double res = e - E0;
    T = T - res / dEdT;
    double err = fabs(res / E0);
    if ((T < h.tl) || (T > h.th)) {
      T = Told;
      err = 1e-15;
    }
    if (err < kHelmTol) break;
  }
  if (ret) {
    double e, cv;
    helm_full(h, rho, T, abar, zbar, P, e, sound, cv);
  } else {
    double dPdT;
    helm_given1(h, rho, T, abar, zbar, P, dPdT, false);
  }
  return iter > kHelmMaxIter ? kHelmMaxIter : iter;
}

// helm_rhoP_given: Newton for T(rho, P); same conventions as helm_rhoe_given.
__device__ __forceinline__ int helm_rhoP_given(const HelmDev &h, double rho, double P0,
                                               double abar, double zbar, double &T, double &E,
                                               double &sound, bool ret) {
  const double Told = T;
  int iter;
  sound = 0.0;
  for (iter = 1; iter <= kHelmMaxIter; iter++) {
    double p, dPdT;
    helm_given1(h, rho, T, abar, zbar, p, dPdT, true);
    const double res = p - P0;
    T = T - res / dPdT;
    double err = fabs(res / P0);
    if ((T < h.tl) || (T > h.th)) {
      T = Told;
      err = 1e-15;
    }
    if (err < kHelmTol) break;
  }
  if (ret) {
    double p, cv;
    helm_full(h, rho, T, abar, zbar, p, E, sound, cv);
  } else {
    double dEdT;
    helm_given2(h, rho, T, abar, zbar, E, dEdT, false);
  }
  return iter > kHelmMaxIter ? kHelmMaxIter : iter;
}

}  // namespace phlx
