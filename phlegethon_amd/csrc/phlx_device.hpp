// phlx_device.hpp -- device-side physics of the explicit FV MHD update.
//
// State is structure-of-arrays in HBM with no stored ghost zones: a load
// outside the interior is redirected through the boundary map of each
// axis (periodic wrap or mirror) and the mirrored normal components are
// sign-flipped, which reproduces the reference's halo copy + physical
// boundary fill (source/source.F90:2361-2446, 7532-7956) value for value.
//
// Floating-point expressions keep the reference's evaluation order
// (left-to-right, no FMA contraction: the library is compiled with
// -ffp-contract=off) so the FP64 path agrees with a CPU run of the same
// algorithm to round-off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phlx_helm.hpp"
#include "phlx_mlp.hpp"

namespace phlx {

constexpr double kRgas = 8.31446261815324e7;  // source/source.F90:605

enum { BC_PERIODIC = 0, BC_REFLECTIVE = 1, BC_OUTFLOW = 2, BC_DIODE = 3 };
enum { REC_LIM2ND = 0, REC_PPH = 1, REC_LIM5TH = 2 };
enum { FLUX_HLLC = 0, FLUX_LHLLC = 1, FLUX_HLLD = 2, FLUX_LHLLD = 3 };

// Device-resident step control: dt/time/smax never round-trip to the host.
struct Scalars {
  double time, dt, smax, tmax, cfl;
  unsigned long long spare;
  long long step;
  int active;  // 0 once time >= tmax: the step's kernels return at once
  int flags;   // mixed precision: SC_U_CUR | SC_P_STALE (set by the kernels that ran)
};
// U holds the conserved state of the current step (the mixed path's stage 1
// starts from it instead of rebuilding state0 from the primitives)
constexpr int SC_U_CUR = 1;
// the FP64 primitives, eint and b_cc lag U (a mixed step ran since they were
// last rebuilt)
constexpr int SC_P_STALE = 2;

struct Grid {
  int n1, n2, n3, sd, nv, nas, ip;
  long long N;       // interior cells
  long long nf[3];   // faces per direction
  int bc[6];
  double inv_dx[3];
  double gm, gmm1, igmm1, inv_mu, mach_cutoff, mach_ct;
  int recon, flux, all_mach, wb, grav, etot, edot, fasteos, mhd;
  double *P, *U, *U0, *eint, *bcc;
  // mixed precision: the sweeps' inputs in FP32, [slot][cell] with slots
  // 0..nv-1 the primitives (rho and p as FP64-formed WB deviations), nv..nv+2
  // the cell-centred B, nv+3 / nv+4 gamma_e (deviation) / gamma_c; written by
  // the update (and k_pack_f), read by the mixed flux kernels
  float *Pf;
  double *bf[3], *b0f[3];
  double *F[3];
  float *Ff[3];            // FP32 view of F[d] (mixed precision: fluxes stored as FP32)
  double *emf[3];          // two EMF components per face (see k_flux)
  float *emff[3];          // FP32 view of emf[d] (mixed precision: face EMFs stored as FP32)
  signed char *sgn[3];     // contact-upwind sign per face
  double *cor[3];          // corner (edge) EMFs along x1, x2, x3
  float *corf[3];          // FP32 view of cor[d] (mixed precision)
  double *ecell;           // cell EMF -(v x B), [comp][cell], written by update / init
  float *ecellf;           // FP32 view of ecell in the mixed path (null in FP64 runs)
  const double *eq_rho_cc, *eq_p_cc, *eq_rho_f, *eq_p_f, *gravy, *phi_cc, *phi_f,
      *edot_cc;
  // Helmholtz FASTEOS: carried temperature (warm start of the Newton solve)
  // and [gamma_e | gamma_c] per cell; WB equilibrium gamma_e profiles
  double *temp, *gam;
  const double *eq_ge_cc, *eq_ge_f;
  HelmDev helm;
  // T(rho, e) start: GUESS_WARM (the reference), GUESS_MLP, GUESS_BRACKET,
  // GUESS_PRED;
  // per-block Newton iteration / zone / fallback counters of the guess
  // variant of the update
  MlpDev mlp;
  int eos_guess;
  int eos_count;           // accumulate the counters (the guess variant only)
  double eos_lin_tol;      // mixed precision: > 0 selects the linearised one-evaluation EoS step
  unsigned int *eos_part;
  float *mlp_y;            // network output at each zone's last solve (NaN: none yet)
  // GUESS_PRED: each zone's last T(rho, e) target e (NaN: none) and c_v there
  double *eprev;
  float *cvf;
  // USE_SHOCK_FLATTENING: detector bytes on cells -2 .. n+1 and flags on
  // cells -1 .. n of every axis (x3 only in 3D); sflat 0 = off
  int sflat;
  double eps_sf;
  unsigned char *sf_det, *sf_flag;
  // THERMAL_DIFFUSION_STS: kappa on cells -1 .. n (one ghost layer), the
  // temperature the residual reads (temp with Helmholtz, an own array for
  // the ideal gas), RKL2 registers e_0, e_jm2, Me_0, wall temperatures of
  // FIX_TEMPERATURE_AT_* (NaN: zero-gradient ghost)
  double *kappa, *tdiff, *e0, *ejm2, *me0;
  double fix_t[6];
  double *red;       // per-block partial maxima of the cell-grid reductions
  int nred;          // number of partials (blocks of a cell-grid launch)
  Scalars *sc;
};

__device__ __forceinline__ long long cix(const Grid &g, int i, int j, int k) {
  return ((long long)k * g.n2 + j) * g.n1 + i;
}
__device__ __forceinline__ long long fix(const Grid &g, int d, int i, int j, int k) {
  if (d == 0) return ((long long)k * g.n2 + j) * (g.n1 + 1) + i;
  if (d == 1) return ((long long)k * (g.n2 + 1) + j) * g.n1 + i;
  return ((long long)k * g.n2 + j) * g.n1 + i;
}
// corner arrays: d = edge direction
__device__ __forceinline__ long long kix(const Grid &g, int d, int i, int j, int k) {
  if (d == 2) return ((long long)k * (g.n2 + 1) + j) * (g.n1 + 1) + i;
  if (d == 0) return ((long long)k * (g.n2 + 1) + j) * g.n1 + i;
  return ((long long)k * g.n2 + j) * (g.n1 + 1) + i;
}

// Ye and abar of a cell from the primitive store: the two advected scalars
// (ADVECT_YE_IABAR, nas 2), or 1 and 1 when the run carries none (PIG_EOS,
// whose evaluation ignores them).
__device__ __forceinline__ void cell_ye_abar(const Grid &g, long long c, double &ye,
                                             double &abar) {
  if (g.nas >= 2) {
    ye = g.P[(long long)(g.sd + 2) * g.N + c];
    abar = 1.0 / g.P[(long long)(g.sd + 3) * g.N + c];
  } else {
    ye = 1.0;
    abar = 1.0;
  }
}

// Shock-flattening arrays: flag box of cells -1 .. n, detector box of cells
// -2 .. n+1 on each axis (the x3 extension only in 3D).
__device__ __forceinline__ long long sfix(const Grid &g, int i, int j, int k) {
  const int k0 = (g.sd == 3) ? 1 : 0;
  return ((long long)(k + k0) * (g.n2 + 2) + (j + 1)) * (g.n1 + 2) + (i + 1);
}
__device__ __forceinline__ long long sdix(const Grid &g, int i, int j, int k) {
  const int k0 = (g.sd == 3) ? 2 : 0;
  return ((long long)(k + k0) * (g.n2 + 4) + (j + 2)) * (g.n1 + 4) + (i + 2);
}

// Map a cell-centred index along axis a into the interior.  side = -1/+1
// when a physical (mirror) boundary was crossed, 0 otherwise.
__device__ __forceinline__ int map_axis(const Grid &g, int a, int c, int n, int &side) {
  side = 0;
  if (c < 0) {
    if (g.bc[2 * a] == BC_PERIODIC) return c + n;
    side = -1;
    return -1 - c;
  }
  if (c >= n) {
    if (g.bc[2 * a + 1] == BC_PERIODIC) return c - n;
    side = 1;
    return 2 * n - 1 - c;
  }
  return c;
}

// Does the mirrored ghost flip the normal (axis a) velocity / field?
// REFLECTIVE always; DIODE when the boundary cell flows inward
// (source/source.F90:7556-7569).
__device__ __forceinline__ bool flips(const Grid &g, int a, int side, long long cb) {
  int bc = g.bc[2 * a + (side > 0)];
  if (bc == BC_REFLECTIVE) return true;
  if (bc == BC_DIODE) {
    double v = g.P[(long long)(1 + a) * g.N + cb];
    return side < 0 ? (v > 0.0) : (v < 0.0);
  }
  return false;
}

// Load primitive state (WB deviation when g.wb) and cell-centred B of a
// possibly-ghost cell.
// FAST: the caller has established that (i, j, k) is an interior cell, so
// the boundary map and the mirror flips are skipped (same values).
template <int SD, int NV, bool MHD, bool GAM = false, bool FAST = false>
__device__ __forceinline__ void load_cell(const Grid &g, int i, int j, int k, double *q,
                                          double *b, double *ga = nullptr) {
  int si = 0, sj = 0, sk = 0;
  int mi = FAST ? i : map_axis(g, 0, i, g.n1, si);
  int mj = FAST ? j : map_axis(g, 1, j, g.n2, sj);
  int mk = (SD == 3 && !FAST) ? map_axis(g, 2, k, g.n3, sk) : k;
  long long c = cix(g, mi, mj, mk);
  const long long N = g.N;
#pragma unroll
  for (int v = 0; v < NV; v++) q[v] = g.P[v * N + c];
  if (g.wb) {
    q[0] = q[0] - g.eq_rho_cc[mj];
    q[SD + 1] = q[SD + 1] - g.eq_p_cc[mj];
  }
  if (GAM) {
    // gamma_e (WB deviation, source/source.F90:7160-7162) and gamma_c; the
    // ghost copies carry no sign flip
    ga[0] = g.gam[c];
    ga[1] = g.gam[N + c];
    if (g.wb) ga[0] = ga[0] - g.eq_ge_cc[mj];
  }
  if (MHD) {
#pragma unroll
    for (int v = 0; v < SD; v++) b[v] = g.bcc[v * N + c];
  }
  if (FAST) return;
  if (si && flips(g, 0, si, cix(g, si < 0 ? 0 : g.n1 - 1, mj, mk))) {
    q[1] = -q[1];
    if (MHD) b[0] = -b[0];
  }
  if (sj && flips(g, 1, sj, cix(g, mi, sj < 0 ? 0 : g.n2 - 1, mk))) {
    q[2] = -q[2];
    if (MHD) b[1] = -b[1];
  }
  if (SD == 3 && sk && flips(g, 2, sk, cix(g, mi, mj, sk < 0 ? 0 : g.n3 - 1))) {
    q[3] = -q[3];
    if (MHD) b[2] = -b[2];
  }
}

// flips() on the FP32 copy of the state (the mixed path writes the FP64
// primitives only at the end of a step): the DIODE test reads the boundary
// cell's normal velocity from g.Pf.
__device__ __forceinline__ bool flips_f(const Grid &g, int a, int side, long long cb) {
  int bc = g.bc[2 * a + (side > 0)];
  if (bc == BC_REFLECTIVE) return true;
  if (bc == BC_DIODE) {
    const float v = g.Pf[(long long)(1 + a) * g.N + cb];
    return side < 0 ? (v > 0.0f) : (v < 0.0f);
  }
  return false;
}

// load_cell for the mixed-precision sweeps: the same values from the FP32
// copy g.Pf (the deviations were formed in FP64 before the cast; the mirror
// flips are exact in either precision).
template <int SD, int NV, bool MHD, bool GAM = false, bool FAST = false>
__device__ __forceinline__ void load_cell_f(const Grid &g, int i, int j, int k, float *q,
                                            float *b, float *ga = nullptr) {
  int si = 0, sj = 0, sk = 0;
  int mi = FAST ? i : map_axis(g, 0, i, g.n1, si);
  int mj = FAST ? j : map_axis(g, 1, j, g.n2, sj);
  int mk = (SD == 3 && !FAST) ? map_axis(g, 2, k, g.n3, sk) : k;
  const long long c = cix(g, mi, mj, mk);
  const long long N = g.N;
#pragma unroll
  for (int v = 0; v < NV; v++) q[v] = g.Pf[v * N + c];
  if (GAM) {
    ga[0] = g.Pf[(NV + 3) * N + c];
    ga[1] = g.Pf[(NV + 4) * N + c];
  }
  if (MHD) {
#pragma unroll
    for (int v = 0; v < SD; v++) b[v] = g.Pf[(NV + v) * N + c];
  }
  if (FAST) return;
  if (si && flips_f(g, 0, si, cix(g, si < 0 ? 0 : g.n1 - 1, mj, mk))) {
    q[1] = -q[1];
    if (MHD) b[0] = -b[0];
  }
  if (sj && flips_f(g, 1, sj, cix(g, mi, sj < 0 ? 0 : g.n2 - 1, mk))) {
    q[2] = -q[2];
    if (MHD) b[1] = -b[1];
  }
  if (SD == 3 && sk && flips_f(g, 2, sk, cix(g, mi, mj, sk < 0 ? 0 : g.n3 - 1))) {
    q[3] = -q[3];
    if (MHD) b[2] = -b[2];
  }
}

// ------------------------------------------------------------ reconstruction
// Face values of one cell from (q-, q, q+): l = left edge, r = right edge.
__device__ __forceinline__ void rec_lim2nd(double qm, double qc, double qp, double &l,
                                           double &r) {
  // harmonic-mean limited slope, source/source.F90:8141-8181
  double sf = qp - qc, sb = qc - qm;
  double prod = sf * sb, slope = 0.0;
  if (prod > 0.0) slope = prod / (sf + sb);
  double h = 0.5 * slope;
  l = qc - h;
  r = qc + h;
}
__device__ __forceinline__ void rec_pph(double qm, double qc, double qp, double &l,
                                        double &r) {
  // unlimited parabola through three cell means, source/source.F90:8183-8193
  double c = (5.0 / 6.0) * qc;
  l = (1.0 / 3.0) * qm + c - (1.0 / 6.0) * qp;
  r = -(1.0 / 6.0) * qm + c + (1.0 / 3.0) * qp;
}
__device__ __forceinline__ void rec_pph_lim(double qm, double qc, double qp, double &l,
                                            double &r) {
  // parabola clipped to the neighbour range + extremum limiter for the
  // advected scalars, source/source.F90:8194-8243
  double c = (5.0 / 6.0) * qc;
  double am = (1.0 / 3.0) * qm + c - (1.0 / 6.0) * qp;
  double ap = -(1.0 / 6.0) * qm + c + (1.0 / 3.0) * qp;
  double lo = fmin(qm, qc), hi = fmax(qm, qc);
  am = am < lo ? lo : am;
  am = am > hi ? hi : am;
  lo = fmin(qc, qp);
  hi = fmax(qc, qp);
  ap = ap < lo ? lo : ap;
  ap = ap > hi ? hi : ap;
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  l = am + qc;
  r = ap + qc;
}

__device__ __forceinline__ void rec_lim5(double qmm, double qm, double qc, double qp,
                                         double qpp, double &l, double &r) {
  // 5th-order face values, clipped to the neighbour range, extremum
  // limiter (source/source.F90:8261-8315)
  const double o420 = 1.0 / 420.0;
  const double c = 329.0 * qc;
  double am = (-21.0 * qmm + 189.0 * qm + c - 91.0 * qp + 14.0 * qpp) * o420;
  double ap = (14.0 * qmm - 91.0 * qm + c + 189.0 * qp - 21.0 * qpp) * o420;
  double lo = fmin(qm, qc), hi = fmax(qm, qc);
  am = am < lo ? lo : am;
  am = am > hi ? hi : am;
  lo = fmin(qc, qp);
  hi = fmax(qc, qp);
  ap = ap < lo ? lo : ap;
  ap = ap > hi ? hi : ap;
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0) {
    ap = 0.0;
    am = 0.0;
  } else {
    if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
    if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
  }
  l = am + qc;
  r = ap + qc;
}

// Face states of one face from the cell window w[0..NS-1] (cells f-O ..):
// left = right edge of cell f-1, right = left edge of cell f.
template <bool R5>
__device__ __forceinline__ void face_states(int recon, bool limited_scalar, const double *w,
                                            int stride, double &qL, double &qR) {
  double l, r;
  if (R5) {
    rec_lim5(w[0], w[stride], w[2 * stride], w[3 * stride], w[4 * stride], l, qL);
    rec_lim5(w[stride], w[2 * stride], w[3 * stride], w[4 * stride], w[5 * stride], qR, r);
  } else if (recon == REC_LIM2ND) {
    rec_lim2nd(w[0], w[stride], w[2 * stride], l, qL);
    rec_lim2nd(w[stride], w[2 * stride], w[3 * stride], qR, r);
  } else if (!limited_scalar) {
    rec_pph(w[0], w[stride], w[2 * stride], l, qL);
    rec_pph(w[stride], w[2 * stride], w[3 * stride], qR, r);
  } else {
    rec_pph_lim(w[0], w[stride], w[2 * stride], l, qL);
    rec_pph_lim(w[stride], w[2 * stride], w[3 * stride], qR, r);
  }
}

// As face_states with USE_SHOCK_FLATTENING: a flattened cell (flm: cell
// f-1, flc: cell f) takes the LIM2ND slope whatever the scheme (x1
// source/source.F90:8743-8869, x2 :9635-9761, x3 :10765-10891).
template <bool R5>
__device__ __forceinline__ void face_states_sf(int recon, bool limited_scalar, const double *w,
                                               int stride, bool flm, bool flc, double &qL,
                                               double &qR) {
  // a[0] = cell f-2, a[s] = f-1, a[2s] = f, a[3s] = f+1
  const double *a = w + (R5 ? stride : 0);
  const int s = stride;
  double l, r;
  if (flm)
    rec_lim2nd(a[0], a[s], a[2 * s], l, qL);
  else if (R5)
    rec_lim5(w[0], w[s], w[2 * s], w[3 * s], w[4 * s], l, qL);
  else if (recon == REC_LIM2ND)
    rec_lim2nd(a[0], a[s], a[2 * s], l, qL);
  else if (!limited_scalar)
    rec_pph(a[0], a[s], a[2 * s], l, qL);
  else
    rec_pph_lim(a[0], a[s], a[2 * s], l, qL);
  if (flc)
    rec_lim2nd(a[s], a[2 * s], a[3 * s], qR, r);
  else if (R5)
    rec_lim5(w[s], w[2 * s], w[3 * s], w[4 * s], w[5 * s], qR, r);
  else if (recon == REC_LIM2ND)
    rec_lim2nd(a[s], a[2 * s], a[3 * s], qR, r);
  else if (!limited_scalar)
    rec_pph(a[s], a[2 * s], a[3 * s], qR, r);
  else
    rec_pph_lim(a[s], a[2 * s], a[3 * s], qR, r);
}

// ------------------------------------------------------------ Riemann: hydro
// HLLC / LHLLC in the face frame (slot 1 = normal velocity).
// source/source.F90:15611-16260.  g_e/g_c: FASTEOS gamma_e, gamma_c.
// EoS kind of the face states: EK_IDEAL, EK_FASTEOS (gamma_e / gamma_c face
// values), EK_HELM_FACE (Helmholtz without FASTEOS: a Newton T(rho, P) per
// face state from the face guess Tg).
enum { EK_IDEAL = 0, EK_FASTEOS = 1, EK_HELM_FACE = 2 };

// rho*e and the sound speed of one face state by helm_rhoP_given from Tg
// (HLLC source/source.F90:15776-15780, :15891-15895; HLLD :14172-14176 and
// the others alike); Ye and abar from the advected scalars (:15706-15725).
template <int SD, int NAS>
__device__ __forceinline__ void face_helm(const Grid &g, const double *q, double Tg,
                                          double &rhoe, double &sound) {
  constexpr int IE = SD + 1, IA = SD + 2;
  const double rho = q[0], p = q[IE];
  double ye = 1.0, abar = 1.0;   // PIG_EOS without scalars ignores both
  if constexpr (NAS >= 2) {
    ye = q[IA];
    abar = 1.0 / q[IA + 1];
  }
  const double zbar = ye * abar;
  double T = Tg, eint;
  helm_rhoP_given(g.helm, rho, p, abar, zbar, T, eint, sound, true);
  rhoe = eint * rho;
}

// SFL: the face may take the two-wave HLL flux of USE_SHOCK_FLATTENING
// (hll true: a cell beside the face is flattened).
template <int SD, int NAS, int EK = EK_IDEAL, bool SFL = false>
__device__ __forceinline__ void riemann_hydro(const Grid &g, const double *qL,
                                              const double *qR, const double *gL,
                                              const double *gR, double Tg, double *F,
                                              bool hll = false) {
  constexpr int IE = SD + 1, IA = SD + 2, NV = SD + 2 + NAS;
  const double rhoL = qL[0], irL = 1.0 / rhoL;
  const double uL = qL[1], vL = qL[2], wL = (SD == 3) ? qL[3] : 0.0;
  const double v2L = uL * uL + vL * vL + wL * wL;
  const double pL = qL[IE];
  double eL, cL, eR, cR;
  if (EK == EK_FASTEOS) {
    eL = pL / (gL[0] - 1.0);
    cL = sqrt(gL[1] * pL * irL);
  } else if (EK == EK_HELM_FACE) {
    face_helm<SD, NAS>(g, qL, Tg, eL, cL);
  } else {
    eL = g.igmm1 * pL;
    cL = sqrt(g.gm * pL * irL);
  }
  const double EL = eL + 0.5 * rhoL * v2L;
  const double rhoR = qR[0], irR = 1.0 / rhoR;
  const double uR = qR[1], vR = qR[2], wR = (SD == 3) ? qR[3] : 0.0;
  const double v2R = uR * uR + vR * vR + wR * wR;
  const double pR = qR[IE];
  if (EK == EK_FASTEOS) {
    eR = pR / (gR[0] - 1.0);
    cR = sqrt(gR[1] * pR * irR);
  } else if (EK == EK_HELM_FACE) {
    face_helm<SD, NAS>(g, qR, Tg, eR, cR);
  } else {
    eR = g.igmm1 * pR;
    cR = sqrt(g.gm * pR * irR);
  }
  const double ER = eR + 0.5 * rhoR * v2R;
  const double cmax = fmax(cL, cR);
  const double sL = fmin(uL, uR) - cmax;
  const double sR = fmax(uL, uR) + cmax;
  if (SFL && hll) {
    // HLL flux with the outer speeds clipped at zero (source/source.F90:15929-16003)
    const double s0 = fmin(0.0, sL), s1 = fmax(0.0, sR);
    double UL[NV], UR[NV], fL[NV], fR[NV];
    UL[0] = rhoL; UL[1] = rhoL * uL; UL[2] = rhoL * vL;
    if (SD == 3) UL[3] = rhoL * wL;
    UL[IE] = EL;
    UR[0] = rhoR; UR[1] = rhoR * uR; UR[2] = rhoR * vR;
    if (SD == 3) UR[3] = rhoR * wR;
    UR[IE] = ER;
    fL[0] = rhoL * uL; fL[1] = rhoL * uL * uL + pL; fL[2] = rhoL * uL * vL;
    if (SD == 3) fL[3] = rhoL * uL * wL;
    fL[IE] = (EL + pL) * uL;
    fR[0] = rhoR * uR; fR[1] = rhoR * uR * uR + pR; fR[2] = rhoR * uR * vR;
    if (SD == 3) fR[3] = rhoR * uR * wR;
    fR[IE] = (ER + pR) * uR;
#pragma unroll
    for (int v = IA; v < NV; v++) {
      UL[v] = rhoL * qL[v];
      UR[v] = rhoR * qR[v];
      fL[v] = UL[v] * uL;
      fR[v] = UR[v] * uR;
    }
    const double idn = 1.0 / (s1 - s0), ss = s0 * s1;
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = idn * (s1 * fL[v] - s0 * fR[v] + ss * (UR[v] - UL[v]));
    return;
  }
  const double dL = sL - uL, dR = sR - uR;
  const double us = (pR - pL + rhoL * uL * dL - rhoR * uR * dR) / (rhoL * dL - rhoR * dR);
  const double rsL = rhoL * (dL / (sL - us));
  const double EsL = rsL * (EL * irL + (us - uL) * (us + pL * irL / dL));
  const double rsR = rhoR * (dR / (sR - us));
  const double EsR = rsR * (ER * irR + (us - uR) * (us + pR * irR / dR));

  if (g.flux == FLUX_LHLLC) {
    // low-dissipation pressure, source/source.F90:16014-16030
    double chi = fmin(1.0, fmax(sqrt(v2L) / cL, sqrt(v2R) / cR));
    chi = fmax(g.mach_cutoff, chi);
    const double phi = chi * (2.0 - chi);
    const double rbar = 0.5 * (rhoL + rhoR), cbar = 0.5 * (cL + cR);
    const double ps = 0.5 * (pL + pR) - 0.5 * phi * rbar * cbar * (uR - uL);
    const double s = copysign(1.0, us);
    const double hp = 0.5 * (1.0 + s), hm = 0.5 * (1.0 - s);
    const double rs = hp * rsL + hm * rsR;
    if (g.all_mach && sL > 0.0) {
      F[0] = rhoL * uL;
      F[1] = rhoL * uL * uL + pL;
      F[2] = rhoL * uL * vL;
      if (SD == 3) F[3] = rhoL * uL * wL;
      F[IE] = (EL + pL) * uL;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rhoL * qL[v] * uL;
    } else if (g.all_mach && sR < 0.0) {
      F[0] = rhoR * uR;
      F[1] = rhoR * uR * uR + pR;
      F[2] = rhoR * uR * vR;
      if (SD == 3) F[3] = rhoR * uR * wR;
      F[IE] = (ER + pR) * uR;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rhoR * qR[v] * uR;
    } else {
      F[0] = rs * us;
      F[1] = rs * us * us + ps;
      F[2] = rs * us * (hp * vL + hm * vR);
      if (SD == 3) F[3] = rs * us * (hp * wL + hm * wR);
      F[IE] = (hp * EsL + hm * EsR + ps) * us;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rs * us * (hp * qL[v] + hm * qR[v]);
    }
    return;
  }
  // HLLC, source/source.F90:16124-16252
  const bool left = us > 0.0;
  const double rho = left ? rhoL : rhoR, u = left ? uL : uR, vv = left ? vL : vR,
               ww = left ? wL : wR, E = left ? EL : ER, p = left ? pL : pR;
  const double s = left ? sL : sR, rs = left ? rsL : rsR, Es = left ? EsL : EsR;
  const double *q = left ? qL : qR;
  double f[NV], U[NV];
  U[0] = rho; U[1] = rho * u; U[2] = rho * vv;
  if (SD == 3) U[3] = rho * ww;
  U[IE] = E;
  f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * vv;
  if (SD == 3) f[3] = rho * u * ww;
  f[IE] = (E + p) * u;
#pragma unroll
  for (int v = IA; v < NV; v++) {
    U[v] = rho * q[v];
    f[v] = rho * u * q[v];
  }
  const bool outer = left ? (sL >= 0.0) : (sR < 0.0);
  if (outer) {
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = f[v];
  } else {
    double Us[NV];
    Us[0] = rs; Us[1] = rs * us; Us[2] = rs * vv;
    if (SD == 3) Us[3] = rs * ww;
    Us[IE] = Es;
#pragma unroll
    for (int v = IA; v < NV; v++) Us[v] = rs * q[v];
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = f[v] + s * (Us[v] - U[v]);
  }
}

// -------------------------------------------------------------- Riemann: MHD
// HLLD / LHLLD in the face frame; b = (b_n, b_t1, b_t2).  Returns the
// induction fluxes Fb (Fb[0] = 0).  source/source.F90:13969-15024.
template <int SD, int NAS, int EK = EK_IDEAL, bool SFL = false>
__device__ __forceinline__ void riemann_mhd(const Grid &g, const double *qL,
                                            const double *qR, const double *bL,
                                            const double *bR, const double *gL,
                                            const double *gR, double Tg, double *F,
                                            double *Fb, bool hll = false) {
  constexpr int IE = SD + 1, IA = SD + 2, NV = SD + 2 + NAS;
  const double rhoL = qL[0], irL = 1.0 / rhoL;
  const double uL = qL[1], vL = qL[2], wL = (SD == 3) ? qL[3] : 0.0;
  const double v2L = uL * uL + vL * vL + wL * wL;
  const double pL = qL[IE];
  double e, c2L, c2R;
  if (EK == EK_FASTEOS) {
    e = pL / (gL[0] - 1.0);
    c2L = gL[1] * pL * irL;
  } else if (EK == EK_HELM_FACE) {
    double sound;
    face_helm<SD, NAS>(g, qL, Tg, e, sound);
    c2L = sound * sound;
  } else {
    e = pL * g.igmm1;
    c2L = g.gm * pL * irL;
  }
  double EL = e + 0.5 * rhoL * v2L;
  const double byL = bL[1], bzL = (SD == 3) ? bL[2] : 0.0;
  const double rhoR = qR[0], irR = 1.0 / rhoR;
  const double uR = qR[1], vR = qR[2], wR = (SD == 3) ? qR[3] : 0.0;
  const double v2R = uR * uR + vR * vR + wR * wR;
  const double pR = qR[IE];
  if (EK == EK_FASTEOS) {
    e = pR / (gR[0] - 1.0);
    c2R = gR[1] * pR * irR;
  } else if (EK == EK_HELM_FACE) {
    double sound;
    face_helm<SD, NAS>(g, qR, Tg, e, sound);
    c2R = sound * sound;
  } else {
    e = pR * g.igmm1;
    c2R = g.gm * pR * irR;
  }
  double ER = e + 0.5 * rhoR * v2R;
  const double byR = bR[1], bzR = (SD == 3) ? bR[2] : 0.0;

  const double bx = 0.5 * (bL[0] + bR[0]), bx2 = bx * bx;
  const double pbL = 0.5 * (bx2 + byL * byL + bzL * bzL);
  const double pbR = 0.5 * (bx2 + byR * byR + bzR * bzR);
  EL = EL + pbL;
  ER = ER + pbR;
  const double ptL = pL + pbL, ptR = pR + pbR;
  const double a2L = 2.0 * pbL * irL, a2R = 2.0 * pbR * irR;
  const double an2L = bx2 * irL, an2R = bx2 * irR;
  double t = c2L + a2L;
  const double cL = sqrt(0.5 * (c2L + a2L + sqrt(t * t - 4.0 * c2L * an2L)));
  t = c2R + a2R;
  const double cR = sqrt(0.5 * (c2R + a2R + sqrt(t * t - 4.0 * c2R * an2R)));
  const double cmax = fmax(cL, cR);
  const double sL = fmin(uL, uR) - cmax;
  const double sR = fmax(uL, uR) + cmax;
  if (SFL && hll) {
    // HLL flux with the outer speeds clipped at zero (source/source.F90:14386-14486)
    const double s0 = fmin(0.0, sL), s1 = fmax(0.0, sR);
    double UL[NV], UR[NV], fL[NV], fR[NV];
    UL[0] = rhoL; UL[1] = rhoL * uL; UL[2] = rhoL * vL;
    if (SD == 3) UL[3] = rhoL * wL;
    UL[IE] = EL;
    UR[0] = rhoR; UR[1] = rhoR * uR; UR[2] = rhoR * vR;
    if (SD == 3) UR[3] = rhoR * wR;
    UR[IE] = ER;
    fL[0] = rhoL * uL;
    fL[1] = rhoL * uL * uL + ptL - bx2;
    fL[2] = rhoL * uL * vL - bx * byL;
    if (SD == 3) fL[3] = rhoL * uL * wL - bx * bzL;
    fL[IE] = (EL + ptL) * uL - bx * (bx * uL + byL * vL + bzL * wL);
    fR[0] = rhoR * uR;
    fR[1] = rhoR * uR * uR + ptR - bx2;
    fR[2] = rhoR * uR * vR - bx * byR;
    if (SD == 3) fR[3] = rhoR * uR * wR - bx * bzR;
    fR[IE] = (ER + ptR) * uR - bx * (bx * uR + byR * vR + bzR * wR);
#pragma unroll
    for (int v = IA; v < NV; v++) {
      UL[v] = rhoL * qL[v];
      UR[v] = rhoR * qR[v];
      fL[v] = UL[v] * uL;
      fR[v] = UR[v] * uR;
    }
    const double UbL[3] = {bx, byL, bzL}, UbR[3] = {bx, byR, bzR};
    const double fbL[3] = {0.0, uL * byL - vL * bx, uL * bzL - wL * bx};
    const double fbR[3] = {0.0, uR * byR - vR * bx, uR * bzR - wR * bx};
    const double idn = 1.0 / (s1 - s0), ss = s0 * s1;
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = idn * (s1 * fL[v] - s0 * fR[v] + ss * (UR[v] - UL[v]));
#pragma unroll
    for (int v = 0; v < SD; v++)
      Fb[v] = idn * (s1 * fbL[v] - s0 * fbR[v] + ss * (UbR[v] - UbL[v]));
    return;
  }
  const double dL = sL - uL, dR = sR - uR;
  const double it1 = 1.0 / (dR * rhoR - dL * rhoL);
  const double us = (dR * rhoR * uR - dL * rhoL * uL - ptR + ptL) * it1;
  const double dmL = sL - us, dmR = sR - us;
  double phi = 1.0;
  if (g.flux == FLUX_LHLLD) {
    // source/source.F90:14364-14384
    t = v2L + a2L;
    const double cu2L = 0.5 * (v2L + a2L + sqrt(t * t - 4.0 * v2L * an2L));
    t = v2R + a2R;
    const double cu2R = 0.5 * (v2R + a2R + sqrt(t * t - 4.0 * v2R * an2R));
    const double mL = sqrt(cu2L) / cL, mR = sqrt(cu2R) / cR;
    if (fmax(mL, mR) < 0.6) {
      double chi = fmin(1.0, fmax(mL, mR));
      chi = fmax(g.mach_cutoff, chi);
      phi = chi * (2.0 - chi);
    }
  }
  const double ps = (dR * rhoR * ptL - dL * rhoL * ptR - phi * rhoL * rhoR * dR * dL * (uL - uR)) * it1;
  const double rsL = rhoL * dL / dmL;
  const double rsR = rhoR * dR / dmR;
  const double ia1 = 1.0 / (rhoL * dL * dmL - bx2);
  const double vsL = vL - bx * byL * (us - uL) * ia1;
  const double bysL = byL * (rhoL * dL * dL - bx2) * ia1;
  const double wsL = wL - bx * bzL * (us - uL) * ia1;
  const double bzsL = bzL * (rhoL * dL * dL - bx2) * ia1;
  const double ia2 = 1.0 / (rhoR * dR * dmR - bx2);
  const double vsR = vR - bx * byR * (us - uR) * ia2;
  const double bysR = byR * (rhoR * dR * dR - bx2) * ia2;
  const double wsR = wR - bx * bzR * (us - uR) * ia2;
  const double bzsR = bzR * (rhoR * dR * dR - bx2) * ia2;
  const double vbL = uL * bx + vL * byL + wL * bzL;
  const double vbR = uR * bx + vR * byR + wR * bzR;
  const double vbsL = us * bx + vsL * bysL + wsL * bzsL;
  const double vbsR = us * bx + vsR * bysR + wsR * bzsR;
  const double EsL = (dL * EL - ptL * uL + ps * us + bx * (vbL - vbsL)) / dmL;
  const double EsR = (dR * ER - ptR * uR + ps * us + bx * (vbR - vbsR)) / dmR;
  const double sqL = sqrt(rsL), sqR = sqrt(rsR);
  const double abx = fabs(bx);
  const double ssL = us - abx / sqL;
  const double ssR = us + abx / sqR;
  const double isq = 1.0 / (sqL + sqR);
  const double sg = copysign(1.0, bx);
  const double vss = (sqL * vsL + sqR * vsR + (bysR - bysL) * sg) * isq;
  const double byss = (sqL * bysR + sqR * bysL + sqL * sqR * (vsR - vsL) * sg) * isq;
  const double wss = (sqL * wsL + sqR * wsR + (bzsR - bzsL) * sg) * isq;
  const double bzss = (sqL * bzsR + sqR * bzsL + sqL * sqR * (wsR - wsL) * sg) * isq;
  const double vbss = us * bx + vss * byss + wss * bzss;
  const double EssL = EsL - sqL * (vbsL - vbss) * sg;
  const double EssR = EsR + sqR * (vbsR - vbss) * sg;

  if (g.flux == FLUX_LHLLD) {
    // one state selected by the wave fan, then the physical flux of it
    double rho, u, v, w, E, by, bz, pt;
    if (g.all_mach && sL > 0.0) {
      rho = rhoL; u = uL; v = vL; w = wL; E = EL; by = byL; bz = bzL; pt = ptL;
    } else if (ssL > 0.0) {
      rho = rsL; u = us; v = vsL; w = wsL; E = EsL; by = bysL; bz = bzsL; pt = ps;
    } else if (us > 0.0) {
      rho = rsL; u = us; v = vss; w = wss; E = EssL; by = byss; bz = bzss; pt = ps;
    } else if (ssR > 0.0) {
      rho = rsR; u = us; v = vss; w = wss; E = EssR; by = byss; bz = bzss; pt = ps;
    } else if (!g.all_mach || sR > 0.0) {
      rho = rsR; u = us; v = vsR; w = wsR; E = EsR; by = bysR; bz = bzsR; pt = ps;
    } else {
      rho = rhoR; u = uR; v = vR; w = wR; E = ER; by = byR; bz = bzR; pt = ptR;
    }
    F[0] = rho * u;
    F[1] = rho * u * u + pt - bx2;
    F[2] = rho * u * v - bx * by;
    if (SD == 3) F[3] = rho * u * w - bx * bz;
    F[IE] = (E + pt) * u - bx * (bx * u + v * by + w * bz);
    if (NAS > 0) {
      const double s2 = copysign(1.0, u), hp = 0.5 * (1.0 + s2), hm = 0.5 * (1.0 - s2);
#pragma unroll
      for (int iv = IA; iv < NV; iv++) F[iv] = rho * u * (hp * qL[iv] + hm * qR[iv]);
    }
    Fb[0] = 0.0;
    Fb[1] = u * by - bx * v;
    Fb[2] = u * bz - bx * w;
    return;
  }
  // HLLD: outer flux plus jumps across the outer and Alfven waves
  const bool left = us > 0.0;
  const double rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR,
               w = left ? wL : wR, E = left ? EL : ER, pt = left ? ptL : ptR,
               by = left ? byL : byR, bz = left ? bzL : bzR, vb = left ? vbL : vbR;
  const double *q = left ? qL : qR;
  double f[NV], U[NV], fb[3], Ub[3];
  U[0] = rho; U[1] = rho * u; U[2] = rho * v;
  if (SD == 3) U[3] = rho * w;
  U[IE] = E;
  f[0] = rho * u;
  f[1] = rho * u * u + pt - bx2;
  f[2] = rho * u * v - bx * by;
  if (SD == 3) f[3] = rho * u * w - bx * bz;
  f[IE] = (E + pt) * u - bx * vb;
#pragma unroll
  for (int iv = IA; iv < NV; iv++) {
    U[iv] = rho * q[iv];
    f[iv] = rho * u * q[iv];
  }
  Ub[0] = bx; Ub[1] = by; Ub[2] = bz;
  fb[0] = 0.0;
  fb[1] = u * by - bx * v;
  fb[2] = u * bz - bx * w;
  const double s = left ? sL : sR;
  const bool outer = left ? (sL > 0.0) : (sR < 0.0);
  if (outer) {
#pragma unroll
    for (int iv = 0; iv < NV; iv++) F[iv] = f[iv];
#pragma unroll
    for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv];
    return;
  }
  const double rs = left ? rsL : rsR, vs = left ? vsL : vsR, ws = left ? wsL : wsR,
               Es = left ? EsL : EsR, bys = left ? bysL : bysR, bzs = left ? bzsL : bzsR;
  double Us[NV], Ubs[3];
  Us[0] = rs; Us[1] = rs * us; Us[2] = rs * vs;
  if (SD == 3) Us[3] = rs * ws;
  Us[IE] = Es;
#pragma unroll
  for (int iv = IA; iv < NV; iv++) Us[iv] = rs * q[iv];
  Ubs[0] = bx; Ubs[1] = bys; Ubs[2] = bzs;
  const bool single = left ? (ssL > 0.0) : (ssR < 0.0);
  if (single) {
#pragma unroll
    for (int iv = 0; iv < NV; iv++) F[iv] = f[iv] + s * (Us[iv] - U[iv]);
#pragma unroll
    for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv] + s * (Ubs[iv] - Ub[iv]);
    return;
  }
  const double ss = left ? ssL : ssR, Ess = left ? EssL : EssR;
  double Uss[NV], Ubss[3];
  Uss[0] = rs; Uss[1] = rs * us; Uss[2] = rs * vss;
  if (SD == 3) Uss[3] = rs * wss;
  Uss[IE] = Ess;
#pragma unroll
  for (int iv = IA; iv < NV; iv++) Uss[iv] = rs * q[iv];
  Ubss[0] = bx; Ubss[1] = byss; Ubss[2] = bzss;
#pragma unroll
  for (int iv = 0; iv < NV; iv++) F[iv] = f[iv] + s * (Us[iv] - U[iv]) + ss * (Uss[iv] - Us[iv]);
#pragma unroll
  for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv] + s * (Ubs[iv] - Ub[iv]) + ss * (Ubss[iv] - Ubs[iv]);
}

// ====================================================================
// Mixed precision (PHLX_PREC_MIXED): reconstruction, wave speeds and the
// Riemann fluxes in FP32 on the FP64-formed well-balanced deviations.
//
// The pressure slot of a face state holds the deviation p' = p - p_eq
// (formed in FP64 before the cast); the full pressure p = p_eq + p' is
// passed beside it for the sound speeds and energies.  The normal momentum
// flux is assembled from deviations, so it carries p* - p_eq directly
// (what the reference obtains by subtracting p_eq from the FP64 flux,
// source/source.F90:8953-8959): in FP32 the subtraction of two O(p_eq)
// numbers would leave only ~6e-8 p_eq of precision, far above the
// low-Mach pressure fluctuation M^2 p.  The star pressures are affine in
// the left/right total pressures with weights summing to one, so the same
// formula evaluated on deviations gives p* - p_eq exactly.
// ====================================================================
namespace mx {

// FP32 quotients and square roots on the hardware's reciprocal and square
// root (v_rcp_f32, v_sqrt_f32: 1 ulp).  The IEEE-rounded FP32 division and
// sqrt expand to ~10-instruction sequences each, and the mixed sweeps are
// VALU-issue bound (~25 quotients and 10 roots per face in LIM2ND + HLLD).
__device__ __forceinline__ float qrcp(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float qsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ void rec_lim2nd(float qm, float qc, float qp, float &l, float &r) {
  float sf = qp - qc, sb = qc - qm;
  float prod = sf * sb, slope = 0.0f;
  if (prod > 0.0f) slope = qdiv(prod, sf + sb);
  float h = 0.5f * slope;
  l = qc - h;
  r = qc + h;
}
__device__ __forceinline__ void rec_pph(float qm, float qc, float qp, float &l, float &r) {
  float c = (5.0f / 6.0f) * qc;
  l = (1.0f / 3.0f) * qm + c - (1.0f / 6.0f) * qp;
  r = -(1.0f / 6.0f) * qm + c + (1.0f / 3.0f) * qp;
}
__device__ __forceinline__ void rec_lim5(float qmm, float qm, float qc, float qp, float qpp,
                                         float &l, float &r) {
  const float c = 329.0f * qc;
  float am = (-21.0f * qmm + 189.0f * qm + c - 91.0f * qp + 14.0f * qpp) * (1.0f / 420.0f);
  float ap = (14.0f * qmm - 91.0f * qm + c + 189.0f * qp - 21.0f * qpp) * (1.0f / 420.0f);
  am = fminf(fmaxf(am, fminf(qm, qc)), fmaxf(qm, qc));
  ap = fminf(fmaxf(ap, fminf(qc, qp)), fmaxf(qc, qp));
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0f) {
    ap = 0.0f;
    am = 0.0f;
  } else {
    if (fabsf(ap) >= 2.0f * fabsf(am)) ap = -2.0f * am;
    if (fabsf(am) >= 2.0f * fabsf(ap)) am = -2.0f * ap;
  }
  l = am + qc;
  r = ap + qc;
}
__device__ __forceinline__ void rec_pph_lim(float qm, float qc, float qp, float &l, float &r) {
  float c = (5.0f / 6.0f) * qc;
  float am = (1.0f / 3.0f) * qm + c - (1.0f / 6.0f) * qp;
  float ap = -(1.0f / 6.0f) * qm + c + (1.0f / 3.0f) * qp;
  am = fminf(fmaxf(am, fminf(qm, qc)), fmaxf(qm, qc));
  ap = fminf(fmaxf(ap, fminf(qc, qp)), fmaxf(qc, qp));
  am = am - qc;
  ap = ap - qc;
  if (ap * am >= 0.0f) {
    ap = 0.0f;
    am = 0.0f;
  } else {
    if (fabsf(ap) >= 2.0f * fabsf(am)) ap = -2.0f * am;
    if (fabsf(am) >= 2.0f * fabsf(ap)) am = -2.0f * ap;
  }
  l = am + qc;
  r = ap + qc;
}

template <bool R5>
__device__ __forceinline__ void face_states(int recon, bool limited_scalar, const float *w,
                                            int stride, float &qL, float &qR) {
  float l, r;
  if (R5) {
    rec_lim5(w[0], w[stride], w[2 * stride], w[3 * stride], w[4 * stride], l, qL);
    rec_lim5(w[stride], w[2 * stride], w[3 * stride], w[4 * stride], w[5 * stride], qR, r);
  } else if (recon == REC_LIM2ND) {
    rec_lim2nd(w[0], w[stride], w[2 * stride], l, qL);
    rec_lim2nd(w[stride], w[2 * stride], w[3 * stride], qR, r);
  } else if (!limited_scalar) {
    rec_pph(w[0], w[stride], w[2 * stride], l, qL);
    rec_pph(w[stride], w[2 * stride], w[3 * stride], qR, r);
  } else {
    rec_pph_lim(w[0], w[stride], w[2 * stride], l, qL);
    rec_pph_lim(w[stride], w[2 * stride], w[3 * stride], qR, r);
  }
}

// HLLC / LHLLC, deviation form.  q[IE] = p', pf = full p.
template <int SD, int NAS, bool FE>
__device__ __forceinline__ void riemann_hydro(const Grid &g, const float *qL, const float *qR,
                                              float pfL, float pfR, const float *gL,
                                              const float *gR, float *F) {
  constexpr int IE = SD + 1, IA = SD + 2, NV = SD + 2 + NAS;
  const float gm = (float)g.gm, igmm1 = (float)g.igmm1;
  const float rhoL = qL[0], irL = qrcp(rhoL);
  const float uL = qL[1], vL = qL[2], wL = (SD == 3) ? qL[3] : 0.0f;
  const float v2L = uL * uL + vL * vL + wL * wL;
  const float rhoR = qR[0], irR = qrcp(rhoR);
  const float uR = qR[1], vR = qR[2], wR = (SD == 3) ? qR[3] : 0.0f;
  const float v2R = uR * uR + vR * vR + wR * wR;
  float eL, cL, eR, cR;
  if (FE) {
    eL = qdiv(pfL, gL[0] - 1.0f);
    cL = qsqrt(gL[1] * pfL * irL);
    eR = qdiv(pfR, gR[0] - 1.0f);
    cR = qsqrt(gR[1] * pfR * irR);
  } else {
    eL = igmm1 * pfL;
    cL = qsqrt(gm * pfL * irL);
    eR = igmm1 * pfR;
    cR = qsqrt(gm * pfR * irR);
  }
  const float EL = eL + 0.5f * rhoL * v2L, ER = eR + 0.5f * rhoR * v2R;
  const float dpL = qL[IE], dpR = qR[IE];
  const float cmax = fmaxf(cL, cR);
  const float sL = fminf(uL, uR) - cmax, sR = fmaxf(uL, uR) + cmax;
  const float dL = sL - uL, dR = sR - uR;
  const float us = qdiv(dpR - dpL + rhoL * uL * dL - rhoR * uR * dR, rhoL * dL - rhoR * dR);
  const float rsL = rhoL * qdiv(dL, sL - us);
  const float EsL = rsL * (EL * irL + (us - uL) * (us + qdiv(pfL * irL, dL)));
  const float rsR = rhoR * qdiv(dR, sR - us);
  const float EsR = rsR * (ER * irR + (us - uR) * (us + qdiv(pfR * irR, dR)));
  if (g.flux == FLUX_LHLLC) {
    float chi = fminf(1.0f, fmaxf(qdiv(qsqrt(v2L), cL), qdiv(qsqrt(v2R), cR)));
    chi = fmaxf((float)g.mach_cutoff, chi);
    const float phi = chi * (2.0f - chi);
    const float rbar = 0.5f * (rhoL + rhoR), cbar = 0.5f * (cL + cR);
    const float diss = 0.5f * phi * rbar * cbar * (uR - uL);
    const float psd = 0.5f * (dpL + dpR) - diss;   // p* - p_eq
    const float psf = 0.5f * (pfL + pfR) - diss;   // p*
    const float s = copysignf(1.0f, us);
    const float hp = 0.5f * (1.0f + s), hm = 0.5f * (1.0f - s);
    const float rs = hp * rsL + hm * rsR;
    if (g.all_mach && sL > 0.0f) {
      F[0] = rhoL * uL;
      F[1] = rhoL * uL * uL + dpL;
      F[2] = rhoL * uL * vL;
      if (SD == 3) F[3] = rhoL * uL * wL;
      F[IE] = (EL + pfL) * uL;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rhoL * qL[v] * uL;
    } else if (g.all_mach && sR < 0.0f) {
      F[0] = rhoR * uR;
      F[1] = rhoR * uR * uR + dpR;
      F[2] = rhoR * uR * vR;
      if (SD == 3) F[3] = rhoR * uR * wR;
      F[IE] = (ER + pfR) * uR;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rhoR * qR[v] * uR;
    } else {
      F[0] = rs * us;
      F[1] = rs * us * us + psd;
      F[2] = rs * us * (hp * vL + hm * vR);
      if (SD == 3) F[3] = rs * us * (hp * wL + hm * wR);
      F[IE] = (hp * EsL + hm * EsR + psf) * us;
#pragma unroll
      for (int v = IA; v < NV; v++) F[v] = rs * us * (hp * qL[v] + hm * qR[v]);
    }
    return;
  }
  // HLLC: F = f + s (U* - U).  At low Mach U* - U is a small difference of
  // O(1) states multiplied by s ~ c, so it is formed from the small
  // quantities directly: rho* - rho = rho (u* - u) / (s - u*), and
  // E* - E = (rho* - rho) E / rho + rho* (u* - u) (u* + p / (rho (s - u))).
  const bool left = us > 0.0f;
  const float rho = left ? rhoL : rhoR, u = left ? uL : uR, vv = left ? vL : vR,
              ww = left ? wL : wR, E = left ? EL : ER, pf = left ? pfL : pfR,
              pd = left ? dpL : dpR;
  const float s = left ? sL : sR, rs = left ? rsL : rsR, d = left ? dL : dR;
  const float *q = left ? qL : qR;
  float f[NV];
  f[0] = rho * u; f[1] = rho * u * u + pd; f[2] = rho * u * vv;
  if (SD == 3) f[3] = rho * u * ww;
  f[IE] = (E + pf) * u;
#pragma unroll
  for (int v = IA; v < NV; v++) f[v] = rho * u * q[v];
  const bool outer = left ? (sL >= 0.0f) : (sR < 0.0f);
  if (outer) {
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = f[v];
  } else {
    const float du = us - u;
    const float D = qdiv(rho * du, s - us);
    float dU[NV];
    dU[0] = D;
    dU[1] = D * us + rho * du;
    dU[2] = D * vv;
    if (SD == 3) dU[3] = D * ww;
    dU[IE] = D * qdiv(E, rho) + rs * du * (us + qdiv(pf, rho * d));
#pragma unroll
    for (int v = IA; v < NV; v++) dU[v] = D * q[v];
#pragma unroll
    for (int v = 0; v < NV; v++) F[v] = f[v] + s * dU[v];
  }
}

// HLLD / LHLLD, deviation form (total pressure deviation p' + B^2/2 in
// the momentum flux, full total pressure in the energy flux).
template <int SD, int NAS, bool FE>
__device__ __forceinline__ void riemann_mhd(const Grid &g, const float *qL, const float *qR,
                                            float pfL, float pfR, const float *bL,
                                            const float *bR, const float *gL, const float *gR,
                                            float *F, float *Fb) {
  constexpr int IE = SD + 1, IA = SD + 2, NV = SD + 2 + NAS;
  const float gm = (float)g.gm, igmm1 = (float)g.igmm1;
  const float rhoL = qL[0], irL = qrcp(rhoL);
  const float uL = qL[1], vL = qL[2], wL = (SD == 3) ? qL[3] : 0.0f;
  const float v2L = uL * uL + vL * vL + wL * wL;
  const float rhoR = qR[0], irR = qrcp(rhoR);
  const float uR = qR[1], vR = qR[2], wR = (SD == 3) ? qR[3] : 0.0f;
  const float v2R = uR * uR + vR * vR + wR * wR;
  float eL, eR, c2L, c2R;
  if (FE) {
    eL = qdiv(pfL, gL[0] - 1.0f);
    c2L = gL[1] * pfL * irL;
    eR = qdiv(pfR, gR[0] - 1.0f);
    c2R = gR[1] * pfR * irR;
  } else {
    eL = pfL * igmm1;
    c2L = gm * pfL * irL;
    eR = pfR * igmm1;
    c2R = gm * pfR * irR;
  }
  const float byL = bL[1], bzL = (SD == 3) ? bL[2] : 0.0f;
  const float byR = bR[1], bzR = (SD == 3) ? bR[2] : 0.0f;
  const float bx = 0.5f * (bL[0] + bR[0]), bx2 = bx * bx;
  const float pbL = 0.5f * (bx2 + byL * byL + bzL * bzL);
  const float pbR = 0.5f * (bx2 + byR * byR + bzR * bzR);
  const float EL = eL + 0.5f * rhoL * v2L + pbL;
  const float ER = eR + 0.5f * rhoR * v2R + pbR;
  const float ptL = pfL + pbL, ptR = pfR + pbR;          // full
  const float pdL = qL[IE] + pbL, pdR = qR[IE] + pbR;    // deviation
  const float a2L = 2.0f * pbL * irL, a2R = 2.0f * pbR * irR;
  const float an2L = bx2 * irL, an2R = bx2 * irR;
  float t = c2L + a2L;
  const float cL = qsqrt(0.5f * (c2L + a2L + qsqrt(fmaxf(t * t - 4.0f * c2L * an2L, 0.0f))));
  t = c2R + a2R;
  const float cR = qsqrt(0.5f * (c2R + a2R + qsqrt(fmaxf(t * t - 4.0f * c2R * an2R, 0.0f))));
  const float cmax = fmaxf(cL, cR);
  const float sL = fminf(uL, uR) - cmax, sR = fmaxf(uL, uR) + cmax;
  const float dL = sL - uL, dR = sR - uR;
  const float it1 = qrcp(dR * rhoR - dL * rhoL);
  const float us = (dR * rhoR * uR - dL * rhoL * uL - pdR + pdL) * it1;
  const float dmL = sL - us, dmR = sR - us;
  float phi = 1.0f;
  if (g.flux == FLUX_LHLLD) {
    t = v2L + a2L;
    const float cu2L = 0.5f * (v2L + a2L + qsqrt(fmaxf(t * t - 4.0f * v2L * an2L, 0.0f)));
    t = v2R + a2R;
    const float cu2R = 0.5f * (v2R + a2R + qsqrt(fmaxf(t * t - 4.0f * v2R * an2R, 0.0f)));
    const float mL = qdiv(qsqrt(cu2L), cL), mR = qdiv(qsqrt(cu2R), cR);
    if (fmaxf(mL, mR) < 0.6f) {
      float chi = fminf(1.0f, fmaxf(mL, mR));
      chi = fmaxf((float)g.mach_cutoff, chi);
      phi = chi * (2.0f - chi);
    }
  }
  const float diss = phi * rhoL * rhoR * dR * dL * (uL - uR);
  const float psf = (dR * rhoR * ptL - dL * rhoL * ptR - diss) * it1;
  const float psd = (dR * rhoR * pdL - dL * rhoL * pdR - diss) * it1;
  const float rsL = qdiv(rhoL * dL, dmL), rsR = qdiv(rhoR * dR, dmR);
  const float ia1 = qrcp(rhoL * dL * dmL - bx2);
  const float vsL = vL - bx * byL * (us - uL) * ia1;
  const float bysL = byL * (rhoL * dL * dL - bx2) * ia1;
  const float wsL = wL - bx * bzL * (us - uL) * ia1;
  const float bzsL = bzL * (rhoL * dL * dL - bx2) * ia1;
  const float ia2 = qrcp(rhoR * dR * dmR - bx2);
  const float vsR = vR - bx * byR * (us - uR) * ia2;
  const float bysR = byR * (rhoR * dR * dR - bx2) * ia2;
  const float wsR = wR - bx * bzR * (us - uR) * ia2;
  const float bzsR = bzR * (rhoR * dR * dR - bx2) * ia2;
  const float vbL = uL * bx + vL * byL + wL * bzL;
  const float vbR = uR * bx + vR * byR + wR * bzR;
  const float vbsL = us * bx + vsL * bysL + wsL * bzsL;
  const float vbsR = us * bx + vsR * bysR + wsR * bzsR;
  const float EsL = qdiv(dL * EL - ptL * uL + psf * us + bx * (vbL - vbsL), dmL);
  const float EsR = qdiv(dR * ER - ptR * uR + psf * us + bx * (vbR - vbsR), dmR);
  const float sqL = qsqrt(rsL), sqR = qsqrt(rsR);
  const float abx = fabsf(bx);
  const float ssL = us - qdiv(abx, sqL), ssR = us + qdiv(abx, sqR);
  const float isq = qrcp(sqL + sqR);
  const float sg = copysignf(1.0f, bx);
  const float vss = (sqL * vsL + sqR * vsR + (bysR - bysL) * sg) * isq;
  const float byss = (sqL * bysR + sqR * bysL + sqL * sqR * (vsR - vsL) * sg) * isq;
  const float wss = (sqL * wsL + sqR * wsR + (bzsR - bzsL) * sg) * isq;
  const float bzss = (sqL * bzsR + sqR * bzsL + sqL * sqR * (wsR - wsL) * sg) * isq;
  const float vbss = us * bx + vss * byss + wss * bzss;
  const float EssL = EsL - sqL * (vbsL - vbss) * sg;
  const float EssR = EsR + sqR * (vbsR - vbss) * sg;

  if (g.flux == FLUX_LHLLD) {
    float rho, u, v, w, E, by, bz, ptd, ptf;
    if (g.all_mach && sL > 0.0f) {
      rho = rhoL; u = uL; v = vL; w = wL; E = EL; by = byL; bz = bzL; ptd = pdL; ptf = ptL;
    } else if (ssL > 0.0f) {
      rho = rsL; u = us; v = vsL; w = wsL; E = EsL; by = bysL; bz = bzsL; ptd = psd; ptf = psf;
    } else if (us > 0.0f) {
      rho = rsL; u = us; v = vss; w = wss; E = EssL; by = byss; bz = bzss; ptd = psd; ptf = psf;
    } else if (ssR > 0.0f) {
      rho = rsR; u = us; v = vss; w = wss; E = EssR; by = byss; bz = bzss; ptd = psd; ptf = psf;
    } else if (!g.all_mach || sR > 0.0f) {
      rho = rsR; u = us; v = vsR; w = wsR; E = EsR; by = bysR; bz = bzsR; ptd = psd; ptf = psf;
    } else {
      rho = rhoR; u = uR; v = vR; w = wR; E = ER; by = byR; bz = bzR; ptd = pdR; ptf = ptR;
    }
    F[0] = rho * u;
    F[1] = rho * u * u + ptd - bx2;
    F[2] = rho * u * v - bx * by;
    if (SD == 3) F[3] = rho * u * w - bx * bz;
    F[IE] = (E + ptf) * u - bx * (bx * u + v * by + w * bz);
    if (NAS > 0) {
      const float s2 = copysignf(1.0f, u), hp = 0.5f * (1.0f + s2), hm = 0.5f * (1.0f - s2);
#pragma unroll
      for (int iv = IA; iv < NV; iv++) F[iv] = rho * u * (hp * qL[iv] + hm * qR[iv]);
    }
    Fb[0] = 0.0f;
    Fb[1] = u * by - bx * v;
    Fb[2] = u * bz - bx * w;
    return;
  }
  // HLLD: F = f + s (U* - U) [+ s** (U** - U*)], the differences formed
  // from small quantities (as in the HLLC branch): with du = u* - u and
  // dm = s - u*, rho* - rho = rho du / dm, v* - v = -bx by du ia,
  // b*_y - b_y = b_y rho (s - u) du ia, and
  // E* - E = (du (E + pt) + (p* - pt) u* + bx (vb - vb*)) / dm.
  const bool left = us > 0.0f;
  const float rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR,
              w = left ? wL : wR, E = left ? EL : ER, ptf = left ? ptL : ptR,
              ptd = left ? pdL : pdR, by = left ? byL : byR, bz = left ? bzL : bzR,
              vb = left ? vbL : vbR;
  const float *q = left ? qL : qR;
  float f[NV], fb[3];
  f[0] = rho * u;
  f[1] = rho * u * u + ptd - bx2;
  f[2] = rho * u * v - bx * by;
  if (SD == 3) f[3] = rho * u * w - bx * bz;
  f[IE] = (E + ptf) * u - bx * vb;
#pragma unroll
  for (int iv = IA; iv < NV; iv++) f[iv] = rho * u * q[iv];
  fb[0] = 0.0f;
  fb[1] = u * by - bx * v;
  fb[2] = u * bz - bx * w;
  const float s = left ? sL : sR;
  const bool outer = left ? (sL > 0.0f) : (sR < 0.0f);
  if (outer) {
#pragma unroll
    for (int iv = 0; iv < NV; iv++) F[iv] = f[iv];
#pragma unroll
    for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv];
    return;
  }
  const float rs = left ? rsL : rsR, vs = left ? vsL : vsR, ws = left ? wsL : wsR,
              bys = left ? bysL : bysR, bzs = left ? bzsL : bzsR,
              vbs = left ? vbsL : vbsR, d = left ? dL : dR, dm = left ? dmL : dmR,
              ia = left ? ia1 : ia2;
  const float du = us - u;
  const float D = qdiv(rho * du, dm);
  float dU[NV], dUb[3];
  dU[0] = D;
  dU[1] = D * us + rho * du;
  dU[2] = D * vs + rho * (-bx * by * du * ia);
  if (SD == 3) dU[3] = D * ws + rho * (-bx * bz * du * ia);
  dU[IE] = qdiv(du * (E + ptf) + (psd - ptd) * us + bx * (vb - vbs), dm);
#pragma unroll
  for (int iv = IA; iv < NV; iv++) dU[iv] = D * q[iv];
  dUb[0] = 0.0f;
  dUb[1] = by * rho * d * du * ia;
  dUb[2] = bz * rho * d * du * ia;
  const bool single = left ? (ssL > 0.0f) : (ssR < 0.0f);
  if (single) {
#pragma unroll
    for (int iv = 0; iv < NV; iv++) F[iv] = f[iv] + s * dU[iv];
#pragma unroll
    for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv] + s * dUb[iv];
    return;
  }
  // double-star state: same density and normal velocity as the star state
  const float ss = left ? ssL : ssR, sq = left ? sqL : sqR;
  float dU2[NV], dUb2[3];
  dU2[0] = 0.0f;
  dU2[1] = 0.0f;
  dU2[2] = rs * (vss - vs);
  if (SD == 3) dU2[3] = rs * (wss - ws);
  dU2[IE] = left ? -sq * (vbs - vbss) * sg : sq * (vbs - vbss) * sg;
#pragma unroll
  for (int iv = IA; iv < NV; iv++) dU2[iv] = 0.0f;
  dUb2[0] = 0.0f;
  dUb2[1] = byss - bys;
  dUb2[2] = bzss - bzs;
#pragma unroll
  for (int iv = 0; iv < NV; iv++) F[iv] = f[iv] + s * dU[iv] + ss * dU2[iv];
#pragma unroll
  for (int iv = 0; iv < 3; iv++) Fb[iv] = fb[iv] + s * dUb[iv] + ss * dUb2[iv];
}

}  // namespace mx

}  // namespace phlx
