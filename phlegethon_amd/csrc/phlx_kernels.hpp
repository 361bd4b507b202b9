// phlx_kernels.hpp -- CDNA4 kernels of the explicit FV MHD update (templates;
// instantiated per precision and reconstruction family by phlx_sel_*.hip).
//
// One RK stage is
//   k_flux<x1>, k_flux<x2>, k_flux<x3>   face states -> Riemann fluxes (+EMF, sign)
//   k_corner                             CT-contact corner EMFs      (MHD)
//   k_faceb                              face-B update (FP64)         (MHD)
//   k_update                             flux divergence + sources + RK combine
//                                        + cons->prim (+ CFL max at the last stage)
// or, fused (k_sweep, 3D with the 2nd-order reconstructions),
//   k_sweep                              the three sweeps + divergence + sources
//                                        + RK combine in one pass over LDS tiles
//   k_corner, k_faceb                    as above (MHD)
//   k_update<..., SWEPT>                 cons->prim (+ CFL max at the last stage)
// Ghost zones are never stored: loads outside the interior go through the
// per-axis boundary map in phlx_device.hpp.  The stencil is not a
// contraction, so there is no MFMA here; every kernel is HBM/VALU work.
#pragma once

#include "phlx_internal.hpp"

namespace phlx {

// Whether every cell of this face's stencil (cells x-O .. x+NS-O-1 along
// DIR) is interior; a wave whose lanes all say yes takes the unmapped loads.
template <int DIR, int NS, int O>
__device__ __forceinline__ bool stencil_inside(const Grid &g, int i, int j, int k) {
  const int x = (DIR == 0) ? i : (DIR == 1) ? j : k;
  const int n = (DIR == 0) ? g.n1 : (DIR == 1) ? g.n2 : g.n3;
  return x - O >= 0 && x + NS - O - 1 < n;
}

// This thread's face (i, j, k) of a sweep along DIR; false past the end.
// x1 faces (n1 + 1 per row) are numbered along the flattened (i, j) plane of
// each k with 256-thread blocks: a 64-wide mapping of rows of n1 + 1 faces
// would add a wave with a single live lane to every row, and the sweep is
// VALU bound.  x2 / x3 faces use the 2D (64, 4) mapping.
// (i, j) of this thread in a flattened e1 x e2 plane of 256-thread blocks
// (blockDim (64, 4)); false past the plane's end.
__device__ __forceinline__ bool plane_coords(int e1, int e2, int &i, int &j) {
  const int t = blockIdx.x * 256 + threadIdx.y * blockDim.x + threadIdx.x;
  int q = (int)((double)t * (1.0 / (double)e1));
  int r = t - q * e1;
  while (r < 0) { q--; r += e1; }
  while (r >= e1) { q++; r -= e1; }
  i = r;
  j = q;
  return j < e2;
}

template <int DIR>
__device__ __forceinline__ bool face_coords(const Grid &g, int &i, int &j, int &k) {
  k = blockIdx.z;
  if (DIR == 0) return plane_coords(g.n1 + 1, g.n2, i, j);
  i = blockIdx.x * blockDim.x + threadIdx.x;
  j = blockIdx.y * blockDim.y + threadIdx.y;
  return i < g.n1 && j < g.n2 + (DIR == 1);
}

// ---------------------------------------------------------------- flux
// Carried T of a (possibly ghost) cell: ghosts copy the mapped interior
// cell's T (the boundary fill flips no sign of T).
template <bool FAST>
__device__ __forceinline__ double cell_temp(const Grid &g, int i, int j, int k) {
  if (FAST) return g.temp[cix(g, i, j, k)];
  int s;
  const int mi = map_axis(g, 0, i, g.n1, s);
  const int mj = map_axis(g, 1, j, g.n2, s);
  const int mk = (g.sd == 3) ? map_axis(g, 2, k, g.n3, s) : k;
  return g.temp[cix(g, mi, mj, mk)];
}

// Helmholtz without FASTEOS: the face's Newton guess, the mean of the T of
// the two cells beside face (i, j, k) along DIR (source/source.F90:8488-8494,
// :9766-9772, :10908-10914).
template <int DIR, bool FAST>
__device__ __forceinline__ double face_temp_guess(const Grid &g, int i, int j, int k) {
  const double tm = cell_temp<FAST>(g, i - (DIR == 0), j - (DIR == 1), k - (DIR == 2));
  const double tc = cell_temp<FAST>(g, i, j, k);
  return 0.5 * (tm + tc);
}

// The face's solve from its reconstructed states (grid frame; transverse B
// and the FASTEOS gammas reconstructed, the normal B read here): rotation
// into the face frame, WB add-back, Riemann solve, rotation back, EMF and
// contact sign (MHD), WB flux correction, ETOT potential flux, store.
// OUT: the fluxes go to Fo instead of the face array (the fused sweep).
template <int DIR, int SD, int NAS, bool MHD, int EK, bool SFL = false, bool OUT = false>
__device__ __forceinline__ void flux_face(const Grid &g, int i, int j, int k, double *qL,
                                          double *qR, double *bL, double *bR, double *gL,
                                          double *gR, double Tg, bool hll = false,
                                          double *Fo = nullptr) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  const long long f = fix(g, DIR, i, j, k);
  if (MHD) {
    const double bn = g.bf[DIR][f];
    bL[DIR] = bn;
    bR[DIR] = bn;
  }
  // into the face frame: x2 (v1,v2)->(v2,-v1); x3 (v1,v2,v3)->(v3,v1,v2)
  if (DIR == 1) {
    double t = qL[1]; qL[1] = qL[2]; qL[2] = -t;
    t = qR[1]; qR[1] = qR[2]; qR[2] = -t;
    if (MHD) {
      t = bL[0]; bL[0] = bL[1]; bL[1] = -t;
      t = bR[0]; bR[0] = bR[1]; bR[1] = -t;
    }
  } else if (DIR == 2) {
    double t = qL[1]; qL[1] = qL[3]; qL[3] = qL[2]; qL[2] = t;
    t = qR[1]; qR[1] = qR[3]; qR[3] = qR[2]; qR[2] = t;
    if (MHD) {
      t = bL[0]; bL[0] = bL[2]; bL[2] = bL[1]; bL[1] = t;
      t = bR[0]; bR[0] = bR[2]; bR[2] = bR[1]; bR[1] = t;
    }
  }
  const int jy = j;  // x2 index of this face (x2-face index for x2, cell index otherwise)
  if (g.wb) {
    const double req = (DIR == 1) ? g.eq_rho_f[jy] : g.eq_rho_cc[jy];
    const double peq = (DIR == 1) ? g.eq_p_f[jy] : g.eq_p_cc[jy];
    qL[0] = qL[0] + req;
    qR[0] = qR[0] + req;
    qL[IE] = qL[IE] + peq;
    qR[IE] = qR[IE] + peq;
  }
  // FASTEOS: WB add-back of the face gamma_e (source/source.F90:8715-8727)
  if (EK == EK_FASTEOS && g.wb) {
    const double geq = (DIR == 1) ? g.eq_ge_f[jy] : g.eq_ge_cc[jy];
    gL[0] = gL[0] + geq;
    gR[0] = gR[0] + geq;
  }
  double F[NV], Fb[3] = {0.0, 0.0, 0.0};
  if (MHD)
    riemann_mhd<SD, NAS, EK, SFL>(g, qL, qR, bL, bR, gL, gR, Tg, F, Fb, hll);
  else
    riemann_hydro<SD, NAS, EK, SFL>(g, qL, qR, gL, gR, Tg, F, hll);
  // back to the grid frame
  if (DIR == 1) {
    double t = F[1]; F[1] = -F[2]; F[2] = t;
    if (MHD) { t = Fb[0]; Fb[0] = -Fb[1]; Fb[1] = t; }
  } else if (DIR == 2) {
    double t = F[1]; F[1] = F[2]; F[2] = F[3]; F[3] = t;
    if (MHD) { t = Fb[0]; Fb[0] = Fb[1]; Fb[1] = Fb[2]; Fb[2] = t; }
  }
  const long long nf = g.nf[DIR];
  if (MHD) {
    // E = -v x B from the induction fluxes (x1 :8926-8932, x2 :10113-10120,
    // x3 :11255-11262); slots: x1 (E2,E3), x2 (E1,E3), x3 (E1,E2)
    double e0, e1v;
    if (DIR == 0) { e0 = Fb[2]; e1v = -Fb[1]; }
    else if (DIR == 1) { e0 = -Fb[2]; e1v = Fb[0]; }
    else { e0 = Fb[1]; e1v = -Fb[0]; }
    g.emf[DIR][f] = e0;
    g.emf[DIR][nf + f] = e1v;
    const double den = 0.5 * (qL[0] + qR[0]) * g.sc->smax;
    const double mach = F[0] / den;
    signed char s = 0;
    if (mach > g.mach_ct) s = 1;
    if (mach < -g.mach_ct) s = -1;
    g.sgn[DIR][f] = s;
  }
  if (g.wb) {
    const double peq = (DIR == 1) ? g.eq_p_f[jy] : g.eq_p_cc[jy];
    F[1 + DIR] = F[1 + DIR] - peq;
  }
  if (g.etot) {
    const double phi = (DIR == 1) ? g.phi_f[jy] : g.phi_cc[jy];
    F[IE] = F[IE] + F[0] * phi;
  }
  if constexpr (OUT) {
#pragma unroll
    for (int v = 0; v < NV; v++) Fo[v] = F[v];
  } else {
#pragma unroll
    for (int v = 0; v < NV; v++) g.F[DIR][v * nf + f] = F[v];
  }
}

// One thread per face along DIR.  Loads the four cells f-2..f+1, builds the
// left state (right edge of cell f-1) and right state (left edge of cell
// f), rotates into the face frame, solves, rotates back and applies the WB
// flux correction and the ETOT potential flux.  MHD faces also store the
// two tangential EMF components and the contact-upwind sign used by CT.
// EK: EoS kind of the face states (EK_IDEAL, EK_FASTEOS, EK_HELM_FACE).
// SF: USE_SHOCK_FLATTENING -- the flags of cells f-1 and f (k_sf_dilate)
// select LIM2ND for a flattened cell and the HLL flux for the face.
template <int DIR, int SD, int NAS, bool MHD, int EK, bool R5, bool SF = false>
__global__ void __launch_bounds__(256) k_flux(Grid g) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  constexpr int NS = R5 ? 6 : 4, O = R5 ? 3 : 2;  // window: cells f-O .. f+NS-O-1
  constexpr bool GAM = EK == EK_FASTEOS;
  if (!g.sc->active) return;
  int i, j, k;
  if (!face_coords<DIR>(g, i, j, k)) return;

  double q[NS][NV], b[NS][3], ga[NS][2];
  double Tg = 0.0;
  if (__all(stencil_inside<DIR, NS, O>(g, i, j, k))) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int o = s - O;
      load_cell<SD, NV, MHD, GAM, true>(g, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0),
                                        k + (DIR == 2 ? o : 0), q[s], b[s], ga[s]);
    }
    if (EK == EK_HELM_FACE)
      Tg = face_temp_guess<DIR, true>(g, i, j, k);
  } else {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int o = s - O;
      load_cell<SD, NV, MHD, GAM>(g, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0),
                                  k + (DIR == 2 ? o : 0), q[s], b[s], ga[s]);
    }
    if (EK == EK_HELM_FACE)
      Tg = face_temp_guess<DIR, false>(g, i, j, k);
  }
  double qL[NV], qR[NV], bL[3] = {0.0, 0.0, 0.0}, bR[3] = {0.0, 0.0, 0.0};
  double gL[2] = {0.0, 0.0}, gR[2] = {0.0, 0.0};
  if (SF) {
    const bool flm = g.sf_flag[sfix(g, i - (DIR == 0), j - (DIR == 1), k - (DIR == 2))] != 0;
    const bool flc = g.sf_flag[sfix(g, i, j, k)] != 0;
#pragma unroll
    for (int v = 0; v < NV; v++)
      face_states_sf<R5>(g.recon, v > IE, &q[0][v], NV, flm, flc, qL[v], qR[v]);
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++) {
        if (c == DIR) continue;
        face_states_sf<R5>(g.recon, false, &b[0][c], 3, flm, flc, bL[c], bR[c]);
      }
    }
    if (GAM) {
#pragma unroll
      for (int c = 0; c < 2; c++)
        face_states_sf<R5>(g.recon, false, &ga[0][c], 2, flm, flc, gL[c], gR[c]);
    }
    // a face beside a flattened cell takes the HLL flux
    // (source/source.F90:8731-8740, :9626-9630, :10756-10760)
    flux_face<DIR, SD, NAS, MHD, EK, true>(g, i, j, k, qL, qR, bL, bR, gL, gR, Tg, flm || flc);
    return;
  }
  // scalars take the limited parabola under PPH (source/source.F90:8194-8243)
#pragma unroll
  for (int v = 0; v < NV; v++) face_states<R5>(g.recon, v > IE, &q[0][v], NV, qL[v], qR[v]);
  if (MHD) {
#pragma unroll
    for (int c = 0; c < SD; c++) {
      if (c == DIR) continue;
      face_states<R5>(g.recon, false, &b[0][c], 3, bL[c], bR[c]);
    }
  }
  // FASTEOS: gamma_e / gamma_c face values, unlimited PPH or the LIM2ND
  // slope (source/source.F90:8497-8713)
  if (GAM) {
#pragma unroll
    for (int c = 0; c < 2; c++) face_states<R5>(g.recon, false, &ga[0][c], 2, gL[c], gR[c]);
  }
  flux_face<DIR, SD, NAS, MHD, EK>(g, i, j, k, qL, qR, bL, bR, gL, gR, Tg);
}

// Both edges of one cell from its window (qm, q0, qp) under the 2nd-order
// reconstructions (LIM2ND, PPH; PPH limits the parabola of scalars).
__device__ __forceinline__ void cell_edges(int recon, bool limited_scalar, double qm, double q0,
                                           double qp, double &l, double &r) {
  if (recon == REC_LIM2ND)
    rec_lim2nd(qm, q0, qp, l, r);
  else if (!limited_scalar)
    rec_pph(qm, q0, qp, l, r);
  else
    rec_pph_lim(qm, q0, qp, l, r);
}

struct CellW {
  double q[8], b[3], ga[2], t;
  int fl;  // shock-flattening flag (SF walks)
};

// x2 / x3 sweeps with the 2nd-order reconstructions: one thread per
// (x1, transverse) column walks kSegFaces consecutive faces along DIR,
// loading each cell once and reconstructing it once; the right edges of
// cell f-1 are carried to face f.  k_flux builds every face from four cell
// loads and two reconstructions.  Each face's arithmetic is that of k_flux.
#ifndef PHLX_SEG_FACES
#define PHLX_SEG_FACES 32
#endif
constexpr int kSegFaces = PHLX_SEG_FACES;
#ifndef PHLX_FLUX_SEG
#define PHLX_FLUX_SEG 1
#endif
// the mixed-precision sweeps' own choice (their faces cost fewer
// instructions, which shifts the balance towards more, shorter threads)
#ifndef PHLX_FLUX_SEG_MX
#define PHLX_FLUX_SEG_MX PHLX_FLUX_SEG
#endif
#ifndef PHLX_SEG_FACES_MX
#define PHLX_SEG_FACES_MX PHLX_SEG_FACES
#endif
constexpr int kSegFacesMx = PHLX_SEG_FACES_MX;

// SF: USE_SHOCK_FLATTENING -- a flagged cell takes the LIM2ND edges, a face
// beside one the HLL flux (as k_flux<..., SF>).
template <int DIR, int SD, int NAS, bool MHD, int EK, bool FAST, bool SF = false>
__device__ __forceinline__ void seg_walk(const Grid &g, int i, int o, int f0, int f1) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  constexpr bool HELM = EK == EK_FASTEOS;  // gamma_e / gamma_c carried per cell
  static_assert(NV <= 8, "cell window size");
  auto load = [&](int c, CellW &w, int ii, int oo) {
    const int jj = DIR == 1 ? c : oo, kk = DIR == 2 ? c : oo;
    load_cell<SD, NV, MHD, HELM, FAST>(g, ii, jj, kk, w.q, w.b, w.ga);
    if (EK == EK_HELM_FACE) w.t = cell_temp<FAST>(g, ii, jj, kk);
    if (SF) {
      // the walk also loads cells -2 and n+1, whose flags it never uses;
      // the flag box holds cells -1 .. n, so their reads are clamped into it
      const int n = DIR == 1 ? g.n2 : g.n3;
      const int cf = min(max(c, -1), n);
      w.fl = g.sf_flag[sfix(g, ii, DIR == 1 ? cf : oo, DIR == 2 ? cf : oo)];
    }
  };
  CellW cm, c0, cp;
  load(f0 - 2, cm, i, o);
  load(f0 - 1, c0, i, o);
  load(f0, cp, i, o);
  // right edges of cell f0 - 1
  double rq[NV], rb[3] = {0.0, 0.0, 0.0}, rg[2] = {0.0, 0.0};
  {
    double l;
    const int rec = (SF && c0.fl) ? REC_LIM2ND : g.recon;
#pragma unroll
    for (int v = 0; v < NV; v++) cell_edges(rec, v > IE, cm.q[v], c0.q[v], cp.q[v], l, rq[v]);
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++)
        if (c != DIR) cell_edges(rec, false, cm.b[c], c0.b[c], cp.b[c], l, rb[c]);
    }
    if (HELM) {
#pragma unroll
      for (int c = 0; c < 2; c++) cell_edges(rec, false, cm.ga[c], c0.ga[c], cp.ga[c], l, rg[c]);
    }
  }
  cm = c0;
  c0 = cp;
#pragma unroll 1
  for (int f = f0; f < f1; f++) {
    // the column indices are re-materialised per face: hoisting every
    // address and profile value of the body out of the loop costs ~100 VGPRs
    int ii = i, oo = o;
    asm volatile("" : "+v"(ii), "+v"(oo));
    load(f + 1, cp, ii, oo);
    double qL[NV], qR[NV], bL[3] = {0.0, 0.0, 0.0}, bR[3] = {0.0, 0.0, 0.0};
    double gL[2] = {0.0, 0.0}, gR[2] = {0.0, 0.0};
    const int rec = (SF && c0.fl) ? REC_LIM2ND : g.recon;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      qL[v] = rq[v];
      cell_edges(rec, v > IE, cm.q[v], c0.q[v], cp.q[v], qR[v], rq[v]);
    }
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++) {
        if (c == DIR) continue;
        bL[c] = rb[c];
        cell_edges(rec, false, cm.b[c], c0.b[c], cp.b[c], bR[c], rb[c]);
      }
    }
    if (HELM) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        gL[c] = rg[c];
        cell_edges(rec, false, cm.ga[c], c0.ga[c], cp.ga[c], gR[c], rg[c]);
      }
    }
    const double Tg = (EK == EK_HELM_FACE) ? 0.5 * (cm.t + c0.t) : 0.0;
    flux_face<DIR, SD, NAS, MHD, EK, SF>(g, ii, DIR == 1 ? f : oo, DIR == 2 ? f : oo, qL, qR, bL,
                                         bR, gL, gR, Tg, SF && (cm.fl || c0.fl));
    cm = c0;
    c0 = cp;
  }
}

template <int DIR, int SD, int NAS, bool MHD, int EK, bool SF = false>
__global__ void __launch_bounds__(256) k_flux_seg(Grid g) {
  static_assert(DIR == 1 || DIR == 2, "x2 / x3 sweeps");
  if (!g.sc->active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int o = blockIdx.y * blockDim.y + threadIdx.y;  // x3 cell (DIR 1) / x2 cell (DIR 2)
  const int n = (DIR == 1) ? g.n2 : g.n3;                // cells along DIR
  if (i >= g.n1 || o >= ((DIR == 1) ? g.n3 : g.n2)) return;
  const int f0 = blockIdx.z * kSegFaces;
  const int f1 = min(f0 + kSegFaces, n + 1);
  // cells f0 - 2 .. f1 are read
  if (__all(f0 - 2 >= 0 && f1 <= n - 1))
    seg_walk<DIR, SD, NAS, MHD, EK, true, SF>(g, i, o, f0, f1);
  else
    seg_walk<DIR, SD, NAS, MHD, EK, false, SF>(g, i, o, f0, f1);
}

// ------------------------------------------------------- flux, mixed precision
// As k_flux with the mixed-precision policy: cells are loaded and the WB
// deviations formed in FP64, then reconstruction, wave speeds and the
// Riemann solve run in FP32 (deviation form, phlx_device.hpp); the face
// equilibrium is added back in FP64 before the cast; fluxes are stored as
// FP32 and accumulated in FP64 by k_update.
// The mixed-precision face solve from its FP32 face states (as flux_face).
// OUT: the fluxes go to Fo instead of the face array (the fused sweep).
template <int DIR, int SD, int NAS, bool MHD, bool HELM, bool OUT = false>
__device__ __forceinline__ void flux_face_mx(const Grid &g, int i, int j, int k, float *qL,
                                             float *qR, float *bL, float *bR, float *gL,
                                             float *gR, float *Fo = nullptr) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  const long long f = fix(g, DIR, i, j, k);
  if (MHD) {
    const float bn = (float)g.bf[DIR][f];
    bL[DIR] = bn;
    bR[DIR] = bn;
  }
  if (DIR == 1) {
    float t = qL[1]; qL[1] = qL[2]; qL[2] = -t;
    t = qR[1]; qR[1] = qR[2]; qR[2] = -t;
    if (MHD) {
      t = bL[0]; bL[0] = bL[1]; bL[1] = -t;
      t = bR[0]; bR[0] = bR[1]; bR[1] = -t;
    }
  } else if (DIR == 2) {
    float t = qL[1]; qL[1] = qL[3]; qL[3] = qL[2]; qL[2] = t;
    t = qR[1]; qR[1] = qR[3]; qR[3] = qR[2]; qR[2] = t;
    if (MHD) {
      t = bL[0]; bL[0] = bL[2]; bL[2] = bL[1]; bL[1] = t;
      t = bR[0]; bR[0] = bR[2]; bR[2] = bR[1]; bR[1] = t;
    }
  }
  const int jy = j;
  // full face pressures for the wave speeds / energies; the pressure slot
  // keeps the deviation (FP64 add-back of the face equilibrium, then cast)
  float pfL = qL[IE], pfR = qR[IE];
  if (g.wb) {
    const double req = (DIR == 1) ? g.eq_rho_f[jy] : g.eq_rho_cc[jy];
    const double peq = (DIR == 1) ? g.eq_p_f[jy] : g.eq_p_cc[jy];
    qL[0] = (float)(req + (double)qL[0]);
    qR[0] = (float)(req + (double)qR[0]);
    pfL = (float)(peq + (double)qL[IE]);
    pfR = (float)(peq + (double)qR[IE]);
  }
  if (HELM && g.wb) {
    const double geq = (DIR == 1) ? g.eq_ge_f[jy] : g.eq_ge_cc[jy];
    gL[0] = (float)(geq + (double)gL[0]);
    gR[0] = (float)(geq + (double)gR[0]);
  }
  float F[NV], Fb[3] = {0.0f, 0.0f, 0.0f};
  if (MHD)
    mx::riemann_mhd<SD, NAS, HELM>(g, qL, qR, pfL, pfR, bL, bR, gL, gR, F, Fb);
  else
    mx::riemann_hydro<SD, NAS, HELM>(g, qL, qR, pfL, pfR, gL, gR, F);
  if (DIR == 1) {
    float t = F[1]; F[1] = -F[2]; F[2] = t;
    if (MHD) { t = Fb[0]; Fb[0] = -Fb[1]; Fb[1] = t; }
  } else if (DIR == 2) {
    float t = F[1]; F[1] = F[2]; F[2] = F[3]; F[3] = t;
    if (MHD) { t = Fb[0]; Fb[0] = Fb[1]; Fb[1] = Fb[2]; Fb[2] = t; }
  }
  const long long nf = g.nf[DIR];
  if (MHD) {
    float e0, e1v;
    if (DIR == 0) { e0 = Fb[2]; e1v = -Fb[1]; }
    else if (DIR == 1) { e0 = -Fb[2]; e1v = Fb[0]; }
    else { e0 = Fb[1]; e1v = -Fb[0]; }
    g.emff[DIR][f] = e0;     // FP32 like the induction fluxes they are
    g.emff[DIR][nf + f] = e1v;
    // |F_rho| / (rho_mean smax) against Mach_ct as a product (den > 0)
    const double den = 0.5 * ((double)qL[0] + (double)qR[0]) * g.sc->smax;
    const double thr = g.mach_ct * den, f0 = (double)F[0];
    signed char sgn = 0;
    if (f0 > thr) sgn = 1;
    if (f0 < -thr) sgn = -1;
    g.sgn[DIR][f] = sgn;
  }
  if (g.etot) {
    const double phi = (DIR == 1) ? g.phi_f[jy] : g.phi_cc[jy];
    F[IE] = (float)((double)F[IE] + (double)F[0] * phi);
  }
  if constexpr (OUT) {
#pragma unroll
    for (int v = 0; v < NV; v++) Fo[v] = F[v];
  } else {
#pragma unroll
    for (int v = 0; v < NV; v++) g.Ff[DIR][v * nf + f] = F[v];
  }
}

// Mixed-precision column segments (as k_flux_seg): cells are loaded and the
// WB deviations formed in FP64, then cast; reconstruction in FP32.
struct CellWf {
  float q[8], b[3], ga[2];
};

__device__ __forceinline__ void cell_edges_mx(int recon, bool limited_scalar, float qm, float q0,
                                              float qp, float &l, float &r) {
  if (recon == REC_LIM2ND)
    mx::rec_lim2nd(qm, q0, qp, l, r);
  else if (!limited_scalar)
    mx::rec_pph(qm, q0, qp, l, r);
  else
    mx::rec_pph_lim(qm, q0, qp, l, r);
}

template <int DIR, int SD, int NAS, bool MHD, bool HELM, bool FAST>
__device__ __forceinline__ void seg_walk_mx(const Grid &g, int i, int o, int f0, int f1) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  static_assert(NV <= 8, "cell window size");
  auto load = [&](int c, CellWf &w, int ii, int oo) {
    w.b[0] = w.b[1] = w.b[2] = 0.0f;
    w.ga[0] = w.ga[1] = 0.0f;
    load_cell_f<SD, NV, MHD, HELM, FAST>(g, ii, DIR == 1 ? c : oo, DIR == 2 ? c : oo, w.q, w.b,
                                         w.ga);
  };
  CellWf cm, c0, cp;
  load(f0 - 2, cm, i, o);
  load(f0 - 1, c0, i, o);
  load(f0, cp, i, o);
  float rq[NV], rb[3] = {0.0f, 0.0f, 0.0f}, rg[2] = {0.0f, 0.0f};
  {
    float l;
#pragma unroll
    for (int v = 0; v < NV; v++)
      cell_edges_mx(g.recon, v > IE, cm.q[v], c0.q[v], cp.q[v], l, rq[v]);
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++)
        if (c != DIR) cell_edges_mx(g.recon, false, cm.b[c], c0.b[c], cp.b[c], l, rb[c]);
    }
    if (HELM) {
#pragma unroll
      for (int c = 0; c < 2; c++)
        cell_edges_mx(g.recon, false, cm.ga[c], c0.ga[c], cp.ga[c], l, rg[c]);
    }
  }
  cm = c0;
  c0 = cp;
  // cell f+1 of the next face is loaded one face ahead, so its latency
  // overlaps a face solve (the walk reads cells f0-2 .. f1 as before)
  CellWf cn;
  load(f0 + 1, cn, i, o);
#pragma unroll 1
  for (int f = f0; f < f1; f++) {
    int ii = i, oo = o;  // re-materialised per face (see seg_walk)
    asm volatile("" : "+v"(ii), "+v"(oo));
    cp = cn;
    if (f + 1 < f1) load(f + 2, cn, ii, oo);
    float qL[NV], qR[NV], bL[3] = {0.0f, 0.0f, 0.0f}, bR[3] = {0.0f, 0.0f, 0.0f};
    float gL[2] = {0.0f, 0.0f}, gR[2] = {0.0f, 0.0f};
#pragma unroll
    for (int v = 0; v < NV; v++) {
      qL[v] = rq[v];
      cell_edges_mx(g.recon, v > IE, cm.q[v], c0.q[v], cp.q[v], qR[v], rq[v]);
    }
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++) {
        if (c == DIR) continue;
        bL[c] = rb[c];
        cell_edges_mx(g.recon, false, cm.b[c], c0.b[c], cp.b[c], bR[c], rb[c]);
      }
    }
    if (HELM) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        gL[c] = rg[c];
        cell_edges_mx(g.recon, false, cm.ga[c], c0.ga[c], cp.ga[c], gR[c], rg[c]);
      }
    }
    flux_face_mx<DIR, SD, NAS, MHD, HELM>(g, ii, DIR == 1 ? f : oo, DIR == 2 ? f : oo, qL, qR,
                                          bL, bR, gL, gR);
    cm = c0;
    c0 = cp;
  }
}

template <int DIR, int SD, int NAS, bool MHD, bool HELM>
__global__ void __launch_bounds__(256) k_flux_seg_mx(Grid g) {
  static_assert(DIR == 1 || DIR == 2, "x2 / x3 sweeps");
  if (!g.sc->active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int o = blockIdx.y * blockDim.y + threadIdx.y;
  const int n = (DIR == 1) ? g.n2 : g.n3;
  if (i >= g.n1 || o >= ((DIR == 1) ? g.n3 : g.n2)) return;
  const int f0 = blockIdx.z * kSegFacesMx;
  const int f1 = min(f0 + kSegFacesMx, n + 1);
  if (__all(f0 - 2 >= 0 && f1 <= n - 1))
    seg_walk_mx<DIR, SD, NAS, MHD, HELM, true>(g, i, o, f0, f1);
  else
    seg_walk_mx<DIR, SD, NAS, MHD, HELM, false>(g, i, o, f0, f1);
}

template <int DIR, int SD, int NAS, bool MHD, bool HELM, bool R5>
__global__ void __launch_bounds__(256) k_flux_mx(Grid g) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  constexpr int NS = R5 ? 6 : 4, O = R5 ? 3 : 2;
  if (!g.sc->active) return;
  int i, j, k;
  if (!face_coords<DIR>(g, i, j, k)) return;

  float q[NS][NV], b[NS][3], ga[NS][2];
  const bool fast = __all(stencil_inside<DIR, NS, O>(g, i, j, k));
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int o = s - O;
    b[s][0] = b[s][1] = b[s][2] = 0.0f;
    ga[s][0] = ga[s][1] = 0.0f;
    if (fast)
      load_cell_f<SD, NV, MHD, HELM, true>(g, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0),
                                           k + (DIR == 2 ? o : 0), q[s], b[s], ga[s]);
    else
      load_cell_f<SD, NV, MHD, HELM>(g, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0),
                                     k + (DIR == 2 ? o : 0), q[s], b[s], ga[s]);
  }
  float qL[NV], qR[NV], bL[3] = {0.0f, 0.0f, 0.0f}, bR[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int v = 0; v < NV; v++) mx::face_states<R5>(g.recon, v > IE, &q[0][v], NV, qL[v], qR[v]);
  if (MHD) {
#pragma unroll
    for (int c = 0; c < SD; c++) {
      if (c == DIR) continue;
      mx::face_states<R5>(g.recon, false, &b[0][c], 3, bL[c], bR[c]);
    }
  }
  float gL[2] = {0.0f, 0.0f}, gR[2] = {0.0f, 0.0f};
  if (HELM) {
#pragma unroll
    for (int c = 0; c < 2; c++) mx::face_states<R5>(g.recon, false, &ga[0][c], 2, gL[c], gR[c]);
  }
  flux_face_mx<DIR, SD, NAS, MHD, HELM>(g, i, j, k, qL, qR, bL, bR, gL, gR);
}

// ------------------------------------------------------- x1, one cell per lane
// x1 sweep with one reconstruction per cell (the column-segment idea in the
// direction the lanes run along): lane l of wave w takes cell
// C = 63 w + l of the plane's row-major sequence of x1 cells with one
// ghost at each end of a row (n1 + 2 per row), reconstructs both of its x1
// edges, and solves the face between its cell and the next lane's (its own
// right edges; the next cell's left edges by a lane shuffle).  Lane 63 and
// the last cell of a row solve nothing; consecutive waves overlap by one
// cell.  Each face's states are those of face_states (the same
// reconstruction calls) and its solve is flux_face / flux_face_mx, so the
// fluxes equal the per-face kernel's bit for bit.  2nd-order
// reconstructions, ideal or FASTEOS states.
#ifndef PHLX_FLUX_X1_CELLS
#define PHLX_FLUX_X1_CELLS 1
#endif
template <int SD, int NAS, bool MHD, int EK, bool MIX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
    k_flux_x1c(Grid g) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  constexpr bool GAM = EK == EK_FASTEOS;
  static_assert(EK != EK_HELM_FACE, "the per-face Newton guess is not shuffled");
  using Re = typename std::conditional<MIX, float, double>::type;
  if (!g.sc->active) return;
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int rowc = g.n1 + 2;
  const long long Cn = (long long)rowc * g.n2;
  const long long C = 63 * wave + lane;
  const int k = blockIdx.y;
  // lanes past the plane stand in for its last cell (no face uses them)
  const long long Cc = C < Cn ? C : Cn - 1;
  const int j = (int)(Cc / rowc), ic = (int)(Cc - (long long)j * rowc) - 1;
  Re q[3][NV], b[3][3], ga[3][2];
  const bool fast = __all(ic - 1 >= 0 && ic + 1 < g.n1);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    b[s][0] = b[s][1] = b[s][2] = 0;
    ga[s][0] = ga[s][1] = 0;
    const int ii = ic - 1 + s;
    if constexpr (MIX) {
      if (fast)
        load_cell_f<SD, NV, MHD, GAM, true>(g, ii, j, k, q[s], b[s], ga[s]);
      else
        load_cell_f<SD, NV, MHD, GAM>(g, ii, j, k, q[s], b[s], ga[s]);
    } else {
      if (fast)
        load_cell<SD, NV, MHD, GAM, true>(g, ii, j, k, q[s], b[s], ga[s]);
      else
        load_cell<SD, NV, MHD, GAM>(g, ii, j, k, q[s], b[s], ga[s]);
    }
  }
  auto edges = [&](bool lim, Re qm, Re q0, Re qp, Re &l, Re &r) {
    if constexpr (MIX)
      cell_edges_mx(g.recon, lim, qm, q0, qp, l, r);
    else
      cell_edges(g.recon, lim, qm, q0, qp, l, r);
  };
  // this cell's edges: right -> the face above (left state), left -> the
  // face below (right state, shuffled to the lane below)
  Re qL[NV], qR[NV], bL[3] = {0, 0, 0}, bR[3] = {0, 0, 0}, gL[2] = {0, 0}, gR[2] = {0, 0};
#pragma unroll
  for (int v = 0; v < NV; v++) {
    Re l;
    edges(v > IE, q[0][v], q[1][v], q[2][v], l, qL[v]);
    qR[v] = __shfl_down(l, 1);
  }
  if (MHD) {
#pragma unroll
    for (int c = 1; c < SD; c++) {
      Re l;
      edges(false, b[0][c], b[1][c], b[2][c], l, bL[c]);
      bR[c] = __shfl_down(l, 1);
    }
  }
  if (GAM) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      Re l;
      edges(false, ga[0][c], ga[1][c], ga[2][c], l, gL[c]);
      gR[c] = __shfl_down(l, 1);
    }
  }
  if (lane == 63 || C >= Cn || ic >= g.n1) return;
  if constexpr (MIX)
    flux_face_mx<0, SD, NAS, MHD, GAM>(g, ic + 1, j, k, qL, qR, bL, bR, gL, gR);
  else
    flux_face<0, SD, NAS, MHD, EK>(g, ic + 1, j, k, qL, qR, bL, bR, gL, gR, 0.0);
}

// ------------------------------------------------------------------ CT
// Face-centred EMF component `comp` (0..2 = E1..E3) of the face array of
// direction fd at a possibly-ghost transverse position; ghosts follow the
// EMF halo copy and EMF boundary conditions (source/source.F90:11370-11964).
// FAST: the caller guarantees an interior position (no map).  F32: the
// mixed path's FP32 face EMFs (emff).
template <bool FAST = false, bool F32 = false>
__device__ __forceinline__ void face_emf(const Grid &g, int fd, int comp, int i, int j,
                                         int k, double &e, double &s) {
  int ii[3] = {i, j, k};
  const int n[3] = {g.n1, g.n2, g.n3};
  double fac = 1.0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (FAST || a == fd || a >= g.sd) continue;
    int c = ii[a];
    if (c < 0 || c >= n[a]) {
      const int side = c < 0 ? 0 : 1;
      if (g.bc[2 * a + side] == BC_PERIODIC) {
        ii[a] = c < 0 ? c + n[a] : c - n[a];
      } else {
        ii[a] = c < 0 ? 0 : n[a] - 1;
        if (g.bc[2 * a + side] == BC_REFLECTIVE && comp != a) fac = -fac;
      }
    }
  }
  const long long f = fix(g, fd, ii[0], ii[1], ii[2]);
  // slot of component comp within the face array
  int slot;
  if (fd == 0) slot = (comp == 1) ? 0 : 1;
  else if (fd == 1) slot = (comp == 0) ? 0 : 1;
  else slot = (comp == 0) ? 0 : 1;
  e = fac * (F32 ? (double)g.emff[fd][slot * g.nf[fd] + f] : g.emf[fd][slot * g.nf[fd] + f]);
  s = (double)g.sgn[fd][f];
}

// -(v x B)_comp of a cell from v and b_cc (the reference forms it from the
// ghost-filled prim and b_cc inside the corner loops, :11966-12185).
template <int SD>
__device__ __forceinline__ double emf_of(int comp, const double *v, const double *b) {
  const double v1 = v[0], v2 = v[1], v3 = (SD == 3) ? v[2] : 0.0;
  if (comp == 2) return -v1 * b[1] + v2 * b[0];
  if (comp == 0) return v3 * b[1] - v2 * b[2];
  return v1 * b[2] - v3 * b[0];
}

// Stored cell EMF of a (possibly ghost) cell.  The update kernel stores
// -(v x B) of every interior cell; a mirrored ghost flips v_a and b_a of the
// mirrored axis a, which negates every component but E_a (each product
// holds exactly one axis-a factor), and IEEE negation commutes with
// rounding, so this equals the value formed from the flipped ghost state
// bit for bit.  F32: the mixed path's FP32 cell EMFs (ecellf).
template <int SD, bool FAST = false, bool F32 = false>
__device__ __forceinline__ double cell_emf(const Grid &g, int comp, int i, int j, int k) {
  if (FAST)
    return F32 ? (double)g.ecellf[(long long)comp * g.N + cix(g, i, j, k)]
               : g.ecell[(long long)comp * g.N + cix(g, i, j, k)];
  int si, sj, sk = 0;
  const int mi = map_axis(g, 0, i, g.n1, si);
  const int mj = map_axis(g, 1, j, g.n2, sj);
  const int mk = (SD == 3) ? map_axis(g, 2, k, g.n3, sk) : k;
  const long long ce = (long long)comp * g.N + cix(g, mi, mj, mk);
  double e = F32 ? (double)g.ecellf[ce] : g.ecell[ce];
  if (comp != 0 && si && flips(g, 0, si, cix(g, si < 0 ? 0 : g.n1 - 1, mj, mk))) e = -e;
  if (comp != 1 && sj && flips(g, 1, sj, cix(g, mi, sj < 0 ? 0 : g.n2 - 1, mk))) e = -e;
  if (SD == 3 && comp != 2 && sk && flips(g, 2, sk, cix(g, mi, mj, sk < 0 ? 0 : g.n3 - 1)))
    e = -e;
  return e;
}

// Gardiner-Stone CT-contact corner EMF (source/source.F90:11966-12027)
__device__ __forceinline__ double corner_emf(double eS, double eW, double eN, double eE,
                                             double eSW, double eNW, double eNE,
                                             double eSE, double sS, double sN, double sW,
                                             double sE) {
  const double dS = 0.5 * (1.0 + sS) * (eW - eSW) + 0.5 * (1.0 - sS) * (eE - eSE);
  const double dN = 0.5 * (1.0 + sN) * (eNW - eW) + 0.5 * (1.0 - sN) * (eNE - eE);
  const double dW = 0.5 * (1.0 + sW) * (eS - eSW) + 0.5 * (1.0 - sW) * (eN - eNW);
  const double dE = 0.5 * (1.0 + sE) * (eSE - eS) + 0.5 * (1.0 - sE) * (eNE - eN);
  return 0.25 * (eS + eW + eN + eE + dS - dN + dW - dE);
}

// The three edge EMFs whose lower end is node (i, j, k) (FAST: all eight
// surrounding cells interior; F32: FP32 face and cell EMFs, FP64 arithmetic).
template <int SD, bool FAST, bool F32>
__device__ __forceinline__ void corner_node(const Grid &g, int i, int j, int k) {
  double eN, eE, eS, eW, sS, sN, sW, sE;
  if (k < g.n3) {  // x3-edges (emfx3_cor)
    face_emf<FAST, F32>(g, 0, 2, i, j, k, eN, sN);
    face_emf<FAST, F32>(g, 1, 2, i, j, k, eE, sE);
    face_emf<FAST, F32>(g, 0, 2, i, j - 1, k, eS, sS);
    face_emf<FAST, F32>(g, 1, 2, i - 1, j, k, eW, sW);
    const double eSW = cell_emf<SD, FAST, F32>(g, 2, i - 1, j - 1, k);
    const double eNW = cell_emf<SD, FAST, F32>(g, 2, i - 1, j, k);
    const double eNE = cell_emf<SD, FAST, F32>(g, 2, i, j, k);
    const double eSE = cell_emf<SD, FAST, F32>(g, 2, i, j - 1, k);
    {
      const double ec = corner_emf(eS, eW, eN, eE, eSW, eNW, eNE, eSE, sS, sN, sW, sE);
      if (F32) g.corf[2][kix(g, 2, i, j, k)] = (float)ec;
      else g.cor[2][kix(g, 2, i, j, k)] = ec;
    }
  }
  if (SD == 3) {
    if (i < g.n1) {  // x1-edges (emfx1_cor, source/source.F90:12050-12107)
      face_emf<FAST, F32>(g, 1, 0, i, j, k, eN, sN);
      face_emf<FAST, F32>(g, 2, 0, i, j, k, eE, sE);
      face_emf<FAST, F32>(g, 1, 0, i, j, k - 1, eS, sS);
      face_emf<FAST, F32>(g, 2, 0, i, j - 1, k, eW, sW);
      const double eSW = cell_emf<SD, FAST, F32>(g, 0, i, j - 1, k - 1);
      const double eNW = cell_emf<SD, FAST, F32>(g, 0, i, j - 1, k);
      const double eNE = cell_emf<SD, FAST, F32>(g, 0, i, j, k);
      const double eSE = cell_emf<SD, FAST, F32>(g, 0, i, j, k - 1);
      {
        const double ec = corner_emf(eS, eW, eN, eE, eSW, eNW, eNE, eSE, sS, sN, sW, sE);
        if (F32) g.corf[0][kix(g, 0, i, j, k)] = (float)ec;
        else g.cor[0][kix(g, 0, i, j, k)] = ec;
      }
    }
    if (j < g.n2) {  // x2-edges (emfx2_cor, source/source.F90:12128-12185)
      face_emf<FAST, F32>(g, 2, 1, i, j, k, eN, sN);
      face_emf<FAST, F32>(g, 0, 1, i, j, k, eE, sE);
      face_emf<FAST, F32>(g, 2, 1, i - 1, j, k, eS, sS);
      face_emf<FAST, F32>(g, 0, 1, i, j, k - 1, eW, sW);
      const double eSW = cell_emf<SD, FAST, F32>(g, 1, i - 1, j, k - 1);
      const double eNW = cell_emf<SD, FAST, F32>(g, 1, i, j, k - 1);
      const double eNE = cell_emf<SD, FAST, F32>(g, 1, i, j, k);
      const double eSE = cell_emf<SD, FAST, F32>(g, 1, i - 1, j, k);
      {
        const double ec = corner_emf(eS, eW, eN, eE, eSW, eNW, eNE, eSE, sS, sN, sW, sE);
        if (F32) g.corf[1][kix(g, 1, i, j, k)] = (float)ec;
        else g.cor[1][kix(g, 1, i, j, k)] = ec;
      }
    }
  }
}

// One thread per grid node (i,j,k) in [0,n1]x[0,n2]x[0,n3]: the three edge
// EMFs whose lower end is that node.  A wave whose nodes all have their
// eight surrounding cells inside the grid reads without the boundary map.
template <int SD, bool F32>
__global__ void __launch_bounds__(256) k_corner(Grid g) {
  if (!g.sc->active) return;
  int i, j;
  const int k = blockIdx.z;
  if (!plane_coords(g.n1 + 1, g.n2 + 1, i, j)) return;
  const bool inside = i >= 1 && i < g.n1 && j >= 1 && j < g.n2 &&
                      (SD == 2 || (k >= 1 && k < g.n3));
  if (__all(inside))
    corner_node<SD, true, F32>(g, i, j, k);
  else
    corner_node<SD, false, F32>(g, i, j, k);
}

// Face-B update b = a1*b0 + a2*b + a3*dt*curl(E) in FP64 (source/source.F90
// :12208-12435).  The update is pointwise in the face field, so where the
// new field goes is free:
//   FACEB_FIRST  b0 = b, kept in b0f; the new field in place of b (RK3);
//   FACEB_LATER  b0 from b0f, b from bf; the new field in place of b;
//   FACEB_RK2_1  b0 = b from bf; the new field into b0f, bf keeps the
//                step-start field (no copy of it); an inactive step copies
//                bf into b0f instead, so b0f holds the current field either
//                way (the host then swaps the two pointer sets);
//   FACEB_RK2_2  b from bf (the stage-1 field, after that swap), b0 from
//                b0f (the step-start field); the new field in place of b0,
//                so after the host swaps back bf holds it.
// F32: the mixed path's FP32 edge EMFs (corf), the update itself in FP64.
template <bool F32>
struct EdgeEmf {
  const double *d;
  const float *f;
  __device__ double operator[](long long x) const { return F32 ? (double)f[x] : d[x]; }
};
__device__ __forceinline__ void faceb_set(const Grid &g, int d, long long f, int mode, double a1,
                                          double a2, double a3, double t) {
  const double b = g.bf[d][f];
  double b0;
  if (mode == FACEB_FIRST) {
    b0 = b;
    g.b0f[d][f] = b;
  } else if (mode == FACEB_RK2_1) {
    b0 = b;
  } else {
    b0 = g.b0f[d][f];
  }
  const double out = a1 * b0 + a2 * b + a3 * t;
  if (mode == FACEB_RK2_1 || mode == FACEB_RK2_2)
    g.b0f[d][f] = out;
  else
    g.bf[d][f] = out;
}
template <int SD, bool F32>
__global__ void __launch_bounds__(256) k_faceb(Grid g, int mode, double a1, double a2,
                                               double c3) {
  int i, j;
  const int k = blockIdx.z;
  if (!plane_coords(g.n1 + 1, g.n2 + 1, i, j)) return;
  if (!g.sc->active) {
    if (mode == FACEB_RK2_1) {
      if (j < g.n2 && k < g.n3) g.b0f[0][fix(g, 0, i, j, k)] = g.bf[0][fix(g, 0, i, j, k)];
      if (i < g.n1 && k < g.n3) g.b0f[1][fix(g, 1, i, j, k)] = g.bf[1][fix(g, 1, i, j, k)];
      if (SD == 3 && i < g.n1 && j < g.n2)
        g.b0f[2][fix(g, 2, i, j, k)] = g.bf[2][fix(g, 2, i, j, k)];
    }
    return;
  }
  const double a3 = c3 * g.sc->dt;
  const double i1 = g.inv_dx[0], i2 = g.inv_dx[1], i3 = g.inv_dx[2];
  const EdgeEmf<F32> C1{g.cor[0], g.corf[0]}, C2{g.cor[1], g.corf[1]}, C3{g.cor[2], g.corf[2]};
  if (j < g.n2 && k < g.n3) {
    double t;
    if (SD == 3)
      t = -(C2[kix(g, 1, i, j, k + 1)] - C2[kix(g, 1, i, j, k)]) * i3 +
          (C3[kix(g, 2, i, j + 1, k)] - C3[kix(g, 2, i, j, k)]) * i2;
    else
      t = (C3[kix(g, 2, i, j + 1, k)] - C3[kix(g, 2, i, j, k)]) * i2;
    faceb_set(g, 0, fix(g, 0, i, j, k), mode, a1, a2, a3, t);
  }
  if (i < g.n1 && k < g.n3) {
    double t;
    if (SD == 3)
      t = (C1[kix(g, 0, i, j, k + 1)] - C1[kix(g, 0, i, j, k)]) * i3 +
          (C3[kix(g, 2, i, j, k)] - C3[kix(g, 2, i + 1, j, k)]) * i1;
    else
      t = (C3[kix(g, 2, i, j, k)] - C3[kix(g, 2, i + 1, j, k)]) * i1;
    faceb_set(g, 1, fix(g, 1, i, j, k), mode, a1, a2, a3, t);
  }
  if (SD == 3 && i < g.n1 && j < g.n2) {
    const double t = (C2[kix(g, 1, i + 1, j, k)] - C2[kix(g, 1, i, j, k)]) * i1 -
                     (C1[kix(g, 0, i, j + 1, k)] - C1[kix(g, 0, i, j, k)]) * i2;
    faceb_set(g, 2, fix(g, 2, i, j, k), mode, a1, a2, a3, t);
  }
}

// ----------------------------------------------------------- wave max
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block maximum written to red[block]; the partials are reduced by one
// small kernel afterwards (a single-address atomic per block serialises
// tens of thousands of blocks at one L2 channel).  max is exact, so the
// result does not depend on the order.
__device__ __forceinline__ void block_max_to(double s, double *red) {
  __shared__ double part[16];
  s = wave_max(s);
  const int lane = threadIdx.x & 63, wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  if (lane == 0) part[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    double m = part[0];
    for (int w = 1; w < nw; w++) m = fmax(m, part[w]);
    red[((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = m;
  }
}

// Reduce n partials with one 1024-thread block.
__device__ __forceinline__ double reduce_partials(const double *red, int n) {
  __shared__ double part[16];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, red[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  m = part[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, part[w]);
  return m;
}

// Signal speed of one cell, summed over directions (source/source.F90
// :13891-13944): (|v_d| + c_f) / dx_d with c_f^2 = c^2 + B^2/rho.
template <int SD, bool MHD, bool HELM, bool KNOWN_CS = false>
__device__ __forceinline__ double cell_speed(const Grid &g, const double *p, const double *b,
                                             double T, double cs_known = 0.0, double ye = 1.0,
                                             double abar = 1.0) {
  double c;
  if (KNOWN_CS) {
    // the sound speed of the zone's own EoS solve (mixed precision)
    c = cs_known;
  } else if (HELM) {
    // full Helmholtz evaluation at the carried T (source/source.F90:13801-13847)
    const double zbar = ye * abar;
    double pr, ei, cv;
    helm_full(g.helm, p[0], T, abar, zbar, pr, ei, c, cv);
  } else {
    c = sqrt(g.gm * p[SD + 1] / p[0]);
  }
  if (MHD) {
    const double bz = (SD == 3) ? b[2] : 0.0;
    const double ca2 = b[0] * b[0] + b[1] * b[1] + bz * bz;
    c = sqrt(c * c + ca2 / p[0]);
  }
  const double s1 = (fabs(p[1]) + c) * g.inv_dx[0];
  const double s2 = (fabs(p[2]) + c) * g.inv_dx[1];
  if (SD == 3) {
    const double s3 = (fabs(p[3]) + c) * g.inv_dx[2];
    return s3 + s1 + s2;
  }
  return s1 + s2;
}

// Block order of a cell kernel launched on a flat grid of nbx * nby * nk
// blocks (x1 tiles of blockDim.x, x2 tiles of blockDim.y, one x3 plane
// each).  PHLX_XCD_KC = 0: the order of the 3D grid (x1 tiles fastest, then
// x2 tiles, then planes).  PHLX_XCD_KC = Kc > 0 (nk % Kc == 0, B % 8 == 0):
// XCD-aware.  Block ids are dealt round-robin to the 8 XCDs (id % 8), each
// with its own L2; XCD x takes the contiguous range of items [x B/8,
// (x+1) B/8), ordered (chunk of Kc planes, tile, plane in chunk), so the x3
// face plane a block reads as its upper face is read next, as the lower face
// of the plane above, by a block on the same XCD.  Bijective:
// id -> (id % 8) B/8 + id / 8, then a mixed-radix decode.
#ifndef PHLX_XCD_KC
#define PHLX_XCD_KC 0
#endif
__device__ __forceinline__ void xcd_cell_block(int nbx, int nby, int nk, int &bx, int &by,
                                               int &bz) {
  const unsigned T = (unsigned)nbx * (unsigned)nby, B = T * (unsigned)nk, id = blockIdx.x;
  constexpr unsigned KC = PHLX_XCD_KC;
  if (KC == 0 || (B & 7u) || ((unsigned)nk % (KC ? KC : 1u))) {
    bx = (int)(id % (unsigned)nbx);
    by = (int)((id / (unsigned)nbx) % (unsigned)nby);
    bz = (int)(id / T);
    return;
  }
  const unsigned w = (id & 7u) * (B >> 3) + (id >> 3);
  const unsigned kk = w % KC, rest = w / KC, t = rest % T;
  bz = (int)((rest / T) * KC + kk);
  bx = (int)(t % (unsigned)nbx);
  by = (int)(t / (unsigned)nbx);
}

// ---------------------------------------------------------------- update
// One thread per cell: state0 rebuild (first stage), flux divergence,
// gravity / heating sources, RK combine, b_cc from the updated faces,
// cons->prim.  At the last stage the cell's CFL signal speed joins the
// device-wide max.
template <bool MIX>
__device__ __forceinline__ double flux_at(const Grid &g, int d, long long idx) {
  return MIX ? (double)g.Ff[d][idx] : g.F[d][idx];
}

// GUESS (Helmholtz only): the T(rho, e) start comes from g.eos_guess (warm,
// surrogate or bracketed) and the Newton iterations / fallbacks are counted
// per block.  The host admits this variant only when every thread of every
// block is an interior cell (nx1 % 64 == 0, nx2 % 4 == 0), so the surrogate's
// wave-wide MFMA and block barriers are reached uniformly.
// MIX with HELM: the T(rho, e) solve is one FP64 evaluation plus a
// linearised Newton step (helm_rhoe_lin) whenever the start's residual is
// small, and that evaluation's sound speed also serves the CFL.
#ifndef PHLX_HELM_UPDATE_WAVES
#define PHLX_HELM_UPDATE_WAVES 1
#endif
// The GUESS variants hold the surrogate's 16 accumulation registers beside
// the EoS state; they are held to 2 waves per SIMD (VGPRs + AGPRs <= 256),
// which the unconstrained allocation can miss by a few registers and then
// halve the occupancy.
// The mixed-precision Helmholtz update (the EoS Newton on a rho-contracted
// cell) is held to PHLX_HELM_MIX_UPDATE_WAVES waves per SIMD: left free, its
// allocation lands a few registers above the 2-wave budget (1 wave per SIMD,
// the table loads and the HBM streaming then overlap poorly).
#ifndef PHLX_HELM_MIX_UPDATE_WAVES
#define PHLX_HELM_MIX_UPDATE_WAVES 2
#endif
template <int SD, int NAS, bool MHD, bool HELM, bool MIX, bool GUESS = false, bool SWEPT = false>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(
        HELM ? (GUESS ? 2 : (MIX ? PHLX_HELM_MIX_UPDATE_WAVES : PHLX_HELM_UPDATE_WAVES)) : 1)))
    k_update(Grid g, int first, int last, double a1,
                                                double a2, double c3, int fin) {
  // first: stage 1, state0 rebuilt from prim + eint (the reference), or in
  // the mixed path taken from U when the device flags say U is current.
  // last: the CFL maximum of the new state.  fin: the step's last stage.
  // The mixed path writes none of the FP64 primitives, eint, B_cc or the
  // gammas (the sweeps read the FP32 copy g.Pf; k_prim_from_u rebuilds the
  // FP64 fields from U and T when something reads them).
  // SWEPT: the fused sweep (k_sweep) has already formed the stage's new
  // conserved state in U; this launch is cons -> prim, EoS and CFL only.
  constexpr int NV = SD + 2 + NAS, IE = SD + 1, IA = SD + 2;
  if (!g.sc->active) return;
  const bool from_u = MIX && first && (g.sc->flags & SC_U_CUR);
  if (MIX && fin && !first && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // a step that ran leaves U current and the FP64 primitives behind it; no
    // thread of this launch reads the flags (first is 0 at the last stage)
    g.sc->flags = SC_U_CUR | SC_P_STALE;
  }
  int bx, by, k;
  xcd_cell_block((g.n1 + blockDim.x - 1) / blockDim.x, (g.n2 + blockDim.y - 1) / blockDim.y, g.n3,
                 bx, by, k);
  const int i = bx * blockDim.x + threadIdx.x;
  const int j = by * blockDim.y + threadIdx.y;
  const bool in = (i < g.n1) && (j < g.n2);
  // The FP64 path saves state0 in U0 and carries the stage state in U.  The
  // mixed path keeps the step's start state in U (no copy at stage 1) and
  // the stage state in U0 until the last stage writes U.
  double *const s0_buf = MIX ? g.U : g.U0;
  double *const s_buf = MIX ? g.U0 : g.U;
  double *const out = (MIX && !fin) ? g.U0 : g.U;
  double smax_local = 0.0;
  if (in) {
    const long long N = g.N, c = cix(g, i, j, k);
    double S[NV], S0[NV], R[NV], p[NV];
    if constexpr (SWEPT) {
#pragma unroll
      for (int v = 0; v < NV; v++) S[v] = out[v * N + c];
    } else {
    // the stored primitives are read only to rebuild state0 (stage 1)
    if (first == 1) {
#pragma unroll
      for (int v = 0; v < NV; v++) p[v] = g.P[v * N + c];
    }
    if (from_u) {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        S0[v] = g.U[v * N + c];
        S[v] = S0[v];
      }
    } else if (first) {
      // source/source.F90:7057-7117
      const double rho = p[0], v1 = p[1], v2 = p[2], v3 = (SD == 3) ? p[3] : 0.0;
      S0[0] = rho;
      S0[1] = rho * v1;
      S0[2] = rho * v2;
      if (SD == 3) S0[3] = rho * v3;
      double E = g.eint[c] + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
      if (MHD) {
        const double b1 = g.bcc[c], b2 = g.bcc[N + c], b3 = (SD == 3) ? g.bcc[2 * N + c] : 0.0;
        E = E + 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
      }
      if (g.etot) E = E + rho * g.phi_cc[j];
      S0[IE] = E;
#pragma unroll
      for (int v = IA; v < NV; v++) S0[v] = rho * p[v];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        S[v] = S0[v];
        s0_buf[v * N + c] = S0[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++) {
        S0[v] = s0_buf[v * N + c];
        S[v] = s_buf[v * N + c];
      }
    }
    // flux divergence: x1 overwrites, x2/x3 accumulate (source/source.F90
    // :9036-9102, :10218-10298, :11322-11359)
    {
      const long long nf = g.nf[0];
      const long long fm = fix(g, 0, i, j, k), fp = fm + 1;
#pragma unroll
      for (int v = 0; v < NV; v++)
        R[v] = (flux_at<MIX>(g, 0, v * nf + fp) - flux_at<MIX>(g, 0, v * nf + fm)) * g.inv_dx[0];
    }
    {
      const long long nf = g.nf[1];
      const long long fm = fix(g, 1, i, j, k), fp = fm + g.n1;
#pragma unroll
      for (int v = 0; v < NV; v++)
        R[v] = R[v] + (flux_at<MIX>(g, 1, v * nf + fp) - flux_at<MIX>(g, 1, v * nf + fm)) * g.inv_dx[1];
    }
    if (SD == 3) {
      const long long nf = g.nf[2];
      const long long fm = fix(g, 2, i, j, k), fp = fm + (long long)g.n1 * g.n2;
#pragma unroll
      for (int v = 0; v < NV; v++)
        R[v] = R[v] + (flux_at<MIX>(g, 2, v * nf + fp) - flux_at<MIX>(g, 2, v * nf + fm)) * g.inv_dx[2];
    }
    if (g.grav) {
      // source/source.F90:12606-12644 on the WB deviation rho' (the stage's
      // prim rho, which is the state's rho bit for bit: cons -> prim stores
      // rho = S[0])
      const double r0 = S[0];
      const double rho = g.wb ? (r0 - g.eq_rho_cc[j]) : r0;
      const double gy = g.gravy[j];
      R[2] = R[2] - rho * gy;
      if (!g.etot) R[IE] = R[IE] - S[2] * gy;
    }
    if (g.edot) R[IE] = R[IE] - g.edot_cc[j];
    const double a3 = c3 * g.sc->dt;
#pragma unroll
    for (int v = 0; v < NV; v++) S[v] = a1 * S0[v] + a2 * S[v] + a3 * R[v];
    }  // !SWEPT
    double b[3] = {0.0, 0.0, 0.0};
    if (MHD) {
      // source/source.F90:12439-12489
      b[0] = 0.5 * (g.bf[0][fix(g, 0, i + 1, j, k)] + g.bf[0][fix(g, 0, i, j, k)]);
      b[1] = 0.5 * (g.bf[1][fix(g, 1, i, j + 1, k)] + g.bf[1][fix(g, 1, i, j, k)]);
      if (SD == 3) b[2] = 0.5 * (g.bf[2][fix(g, 2, i, j, k + 1)] + g.bf[2][fix(g, 2, i, j, k)]);
      if (!MIX) {
#pragma unroll
        for (int v = 0; v < SD; v++) g.bcc[v * N + c] = b[v];
      }
    }
    // cons -> prim (source/source.F90:13533-13686).  Everything but the
    // EoS outputs is stored before the Helmholtz solve, so that little state
    // stays live across it.
    const double rho = S[0], m1 = S[1], m2 = S[2], m3 = (SD == 3) ? S[3] : 0.0;
    const double ir = 1.0 / rho;
    double eint = S[IE] - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * ir;
    if (MHD) eint = eint - 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (g.etot) eint = eint - rho * g.phi_cc[j];
    p[0] = rho;
    p[1] = m1 * ir;
    p[2] = m2 * ir;
    if (SD == 3) p[3] = m3 * ir;
#pragma unroll
    for (int v = IA; v < NV; v++) p[v] = S[v] * ir;
    if (!HELM) p[IE] = g.gmm1 * eint;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (!MIX && (!HELM || v != IE)) g.P[v * N + c] = p[v];
      out[v * N + c] = S[v];
    }
    if (!MIX) g.eint[c] = eint;
    if constexpr (MIX) {
      // FP32 copy of the next sweeps' inputs (the WB deviations formed in FP64)
      g.Pf[c] = (float)(g.wb ? p[0] - g.eq_rho_cc[j] : p[0]);
#pragma unroll
      for (int v = 1; v < NV; v++)
        if (v != IE) g.Pf[v * N + c] = (float)p[v];
      if (!HELM) g.Pf[IE * N + c] = (float)(g.wb ? p[IE] - g.eq_p_cc[j] : p[IE]);
      if (MHD) {
#pragma unroll
        for (int v = 0; v < SD; v++) g.Pf[(NV + v) * N + c] = (float)b[v];
      }
    }
    if (MHD) {
      const double vv[3] = {p[1], p[2], SD == 3 ? p[3] : 0.0};
      if constexpr (MIX) {
        g.ecellf[2 * N + c] = (float)emf_of<SD>(2, vv, b);
        if (SD == 3) {
          g.ecellf[c] = (float)emf_of<SD>(0, vv, b);
          g.ecellf[N + c] = (float)emf_of<SD>(1, vv, b);
        }
      } else {
        g.ecell[2 * N + c] = emf_of<SD>(2, vv, b);
        if (SD == 3) {
          g.ecell[c] = emf_of<SD>(0, vv, b);
          g.ecell[N + c] = emf_of<SD>(1, vv, b);
        }
      }
    }
    double T = 0.0, sound_out = 0.0, ye_c = 1.0, abar_c = 1.0;
    if (HELM) {
      // Ye / 1/abar advected; T(rho, e) Newton from the carried T, then
      // gamma_e = p/eint + 1 and gamma_c = c^2 rho / p (:13573-13578,
      // :13648-13660)
      double ye = 1.0, abar = 1.0;   // PIG_EOS without scalars ignores both
      if constexpr (NAS >= 2) {
        ye = p[IA];
        abar = rho / S[IA + 1];
      }
      if constexpr (NAS >= 2) {
        // the CFL's evaluation takes abar from the stored prim, 1 / (1/abar)
        // (compute_hyperbolic_dt, source/source.F90:13801-13847)
        ye_c = p[IA];
        abar_c = 1.0 / p[IA + 1];
      }
      const double zbar = ye * abar;
      double pr, sound;
      T = g.temp[c];
      const double e0 = eint * ir;
      // GUESS_PRED: start from the energy predictor of the zone's last solve
      const bool pred = g.eos_guess == GUESS_PRED;
      double tg = T, cvl = 0.0;
      if (pred) tg = helm_pred_start(T, e0, g.eprev[c], g.cvf[c]);
      if constexpr (GUESS) {
        __shared__ unsigned int lds[4 * kMlpLdsWaveDw];
        __shared__ unsigned int cnt[3];
        const int mode = g.eos_guess;
        if (mode == GUESS_MLP) {
          // corrected warm start: T_prev * 10^(s (y_now - y_prev)); the
          // previous output is NaN before the first solve of a zone
          float x[4];
          bool ok;
          mlp_inputs(g.mlp, rho, e0, abar, zbar, x, ok);
          const float y = mlp_forward_wave(g.mlp, x, lds + threadIdx.y * kMlpLdsWaveDw);
          const float yp = g.mlp_y[c];
          // 10^d with d ~ the change of log10 T since the last solve:
          // FP32 exp2 (the factor's ~1e-7 error only moves the Newton start)
          if (ok && yp == yp)
            tg = T * (double)__builtin_amdgcn_exp2f(g.mlp.out_scale * (y - yp) *
                                                    3.32192809488736235f);
          g.mlp_y[c] = ok ? y : __int_as_float(0x7fffffff);
        }
        int fb = 0, it;
        if constexpr (MIX) {
          if (mode == GUESS_BRACKET) tg = helm_bracket_guess(g.helm, rho, e0, abar, zbar);
          if (tg >= g.helm.tl && tg <= g.helm.th) T = tg;
          if (g.eos_lin_tol > 0.0)
            it = helm_rhoe_lin(g.helm, g.eos_lin_tol, rho, e0, abar, zbar, T, pr, sound, fb, cvl);
          else
            it = helm_rhoe_mixed(g.helm, rho, e0, abar, zbar, T, pr, sound, cvl);
        } else {
          it = helm_rhoe_guess(g.helm, mode, tg, g.mlp.resid_tol, rho, e0, abar, zbar, T, pr,
                               sound, g.fasteos != 0, fb, &cvl);
        }
        if (g.eos_count) {  // uniform across the grid
          const int tid = threadIdx.y * blockDim.x + threadIdx.x;
          if (tid < 3) cnt[tid] = 0u;
          __syncthreads();
          atomicAdd(&cnt[0], (unsigned int)it);
          atomicAdd(&cnt[1], 1u);
          atomicAdd(&cnt[2], (unsigned int)fb);
          __syncthreads();
          if (tid < 3) {
            const long long blk = blockIdx.x + (long long)gridDim.x *
                                                   (blockIdx.y + (long long)gridDim.y * blockIdx.z);
            g.eos_part[3 * blk + tid] = g.eos_part[3 * blk + tid] + cnt[tid];
          }
        }
      } else if constexpr (MIX) {
        // the reference Newton to its tolerance from the carried T (or the
        // predictor's start), every iterate one FP64 evaluation; a positive
        // eos_lin_tol selects the one-evaluation linearised step instead
        if (tg >= g.helm.tl && tg <= g.helm.th) T = tg;
        int fb;
        if (g.eos_lin_tol > 0.0)
          helm_rhoe_lin(g.helm, g.eos_lin_tol, rho, e0, abar, zbar, T, pr, sound, fb, cvl);
        else
          helm_rhoe_mixed(g.helm, rho, e0, abar, zbar, T, pr, sound, cvl);
      } else {
        // without FASTEOS the last evaluation is pressure only (:13654-13655)
        if (pred && tg >= g.helm.tl && tg <= g.helm.th) T = tg;
        helm_rhoe_given(g.helm, rho, e0, abar, zbar, T, pr, sound, g.fasteos != 0,
                        pred ? &cvl : nullptr);
      }
      if (pred) {
        g.eprev[c] = e0;
        g.cvf[c] = (float)cvl;
      }
      helm_pin(pr);
      helm_pin(sound);
      // the mixed path leaves the FP64 pressure and gammas to k_prim_from_u
      // (an evaluation at the stored T when something reads them)
      if (g.fasteos && !MIX) {
        g.gam[c] = pr / eint + 1.0;
        g.gam[N + c] = sound * sound * rho / pr;
      }
      g.temp[c] = T;
      if (!MIX) g.P[IE * N + c] = pr;
      p[IE] = pr;
      sound_out = sound;
      if constexpr (MIX) {
        g.Pf[IE * N + c] = (float)(g.wb ? pr - g.eq_p_cc[j] : pr);
        const double ge = pr / eint + 1.0;
        g.Pf[(NV + 3) * N + c] = (float)(g.wb ? ge - g.eq_ge_cc[j] : ge);
        g.Pf[(NV + 4) * N + c] = (float)(sound * sound * rho / pr);
      }
    }
    if (last)
      smax_local = cell_speed<SD, MHD, HELM, HELM && MIX>(g, p, b, T, sound_out, ye_c, abar_c);
  }
  if (last) block_max_to(smax_local, g.red);
}

// The mixed path's FP64 primitives, eint and b_cc, rebuilt from the
// conserved state U and the face field by the operations of k_update's
// cons -> prim, so they equal what that kernel would have stored bit for
// bit; with the Helmholtz EoS the pressure and (FASTEOS) gamma_e / gamma_c
// come from one evaluation at the zone's stored T, the result of the last
// stage's Newton solve (source/source.F90:13573-13578; they differ from
// that solve's closing values by the linearisation of its last step, below
// FP64 resolution, and in the sound speed by ~dT/T of that step).  The
// mixed update leaves them to this kernel, which the library runs before
// anything reads them.
template <int SD, int NAS, bool MHD, bool HELM>
__global__ void __launch_bounds__(256) k_prim_from_u(Grid g) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1, IA = SD + 2;
  // only a mixed step that ran leaves the primitives behind U; k_flags_and
  // clears the flag after this launch
  if (!(g.sc->flags & SC_P_STALE)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.n1 || j >= g.n2) return;
  const long long N = g.N, c = cix(g, i, j, k);
  double S[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) S[v] = g.U[v * N + c];
  double b[3] = {0.0, 0.0, 0.0};
  if (MHD) {
    b[0] = 0.5 * (g.bf[0][fix(g, 0, i + 1, j, k)] + g.bf[0][fix(g, 0, i, j, k)]);
    b[1] = 0.5 * (g.bf[1][fix(g, 1, i, j + 1, k)] + g.bf[1][fix(g, 1, i, j, k)]);
    if (SD == 3) b[2] = 0.5 * (g.bf[2][fix(g, 2, i, j, k + 1)] + g.bf[2][fix(g, 2, i, j, k)]);
#pragma unroll
    for (int v = 0; v < SD; v++) g.bcc[v * N + c] = b[v];
  }
  const double rho = S[0], m1 = S[1], m2 = S[2], m3 = (SD == 3) ? S[3] : 0.0;
  const double ir = 1.0 / rho;
  double eint = S[IE] - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * ir;
  if (MHD) eint = eint - 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  if (g.etot) eint = eint - rho * g.phi_cc[j];
  g.P[c] = rho;
  g.P[N + c] = m1 * ir;
  g.P[2 * N + c] = m2 * ir;
  if (SD == 3) g.P[3 * N + c] = m3 * ir;
#pragma unroll
  for (int v = IA; v < NV; v++) g.P[v * N + c] = S[v] * ir;
  if (!HELM) g.P[IE * N + c] = g.gmm1 * eint;
  g.eint[c] = eint;
  if constexpr (HELM) {
    double ye = 1.0, abar = 1.0;  // PIG_EOS without scalars ignores both
    if constexpr (NAS >= 2) {
      ye = S[IA] * ir;
      abar = rho / S[IA + 1];
    }
    double pr, e, sound, cv;
    helm_full(g.helm, rho, g.temp[c], abar, ye * abar, pr, e, sound, cv);
    g.P[IE * N + c] = pr;
    if (g.fasteos) {
      g.gam[c] = pr / eint + 1.0;
      g.gam[N + c] = sound * sound * rho / pr;
    }
  }
}

// Standalone compute_hyperbolic_dt pass over the current prim / b_cc.
template <int SD, bool MHD, bool HELM>
__global__ void __launch_bounds__(256) k_cfl(Grid g) {
  constexpr int NL = SD + 2;  // rho, v, p
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  double s = 0.0;
  if (i < g.n1 && j < g.n2) {
    const long long N = g.N, c = cix(g, i, j, k);
    double p[NL], b[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < NL; v++) p[v] = g.P[v * N + c];
    if (MHD)
#pragma unroll
      for (int v = 0; v < SD; v++) b[v] = g.bcc[v * N + c];
    double ye = 1.0, abar = 1.0;
    if (HELM) cell_ye_abar(g, c, ye, abar);
    s = cell_speed<SD, MHD, HELM>(g, p, b, HELM ? g.temp[c] : 0.0, 0.0, ye, abar);
  }
  block_max_to(s, g.red);
}

// The FP32 copy of the sweep inputs from the FP64 state (mixed precision:
// after init / resume and after a thermal-diffusion step), with the values
// the update writes.
template <int SD, int NAS, bool MHD, bool HELM>
__global__ void __launch_bounds__(256) k_pack_f(Grid g) {
  constexpr int NV = SD + 2 + NAS, IE = SD + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.n1 || j >= g.n2) return;
  const long long N = g.N, c = cix(g, i, j, k);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    double x = g.P[v * N + c];
    if (g.wb && v == 0) x = x - g.eq_rho_cc[j];
    if (g.wb && v == IE) x = x - g.eq_p_cc[j];
    g.Pf[v * N + c] = (float)x;
  }
  if (MHD) {
#pragma unroll
    for (int v = 0; v < SD; v++) g.Pf[(NV + v) * N + c] = (float)g.bcc[v * N + c];
  }
  if (HELM) {
    const double ge = g.gam[c];
    g.Pf[(NV + 3) * N + c] = (float)(g.wb ? ge - g.eq_ge_cc[j] : ge);
    g.Pf[(NV + 4) * N + c] = (float)g.gam[N + c];
  }
}

// Step-0 initialisation: b_cc from faces and the ideal-gas eint
// (source/source.F90:6012-6062, :6373-6376).
template <int SD, bool MHD, bool HELM>
__global__ void __launch_bounds__(256) k_init(Grid g, int restart) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.n1 || j >= g.n2) return;
  const long long N = g.N, c = cix(g, i, j, k);
  if (HELM && restart) {
    // resume: eint (+ gamma_e, gamma_c) from a full evaluation at the
    // restored (rho, T); its pressure only feeds the gammas
    // (source/source.F90:6413-6500)
    const double rho = g.P[c];
    double ye, abar;
    cell_ye_abar(g, c, ye, abar);
    const double zbar = abar * ye;
    double pe, eint, sound, cv;
    helm_full(g.helm, rho, g.temp[c], abar, zbar, pe, eint, sound, cv);
    eint = rho * eint;
    if (g.fasteos) {
      g.gam[c] = pe / eint + 1.0;
      g.gam[N + c] = sound * sound * rho / pe;
    }
    g.eint[c] = eint;
  } else if (HELM) {
    // T(rho, p) Newton from the uploaded T, eint (+ gamma_e, gamma_c with
    // FASTEOS) (source/source.F90:6279-6285, :6351-6361)
    const double rho = g.P[c], p = g.P[(SD + 1) * N + c];
    double ye, abar;
    cell_ye_abar(g, c, ye, abar);
    const double zbar = abar * ye;
    double T = g.temp[c], eint, sound;
    helm_rhoP_given(g.helm, rho, p, abar, zbar, T, eint, sound, g.fasteos != 0);
    eint = rho * eint;
    if (g.fasteos) {
      g.gam[c] = p / eint + 1.0;
      g.gam[N + c] = sound * sound * rho / p;
    }
    g.eint[c] = eint;
    g.temp[c] = T;
  } else {
    g.eint[c] = g.P[(SD + 1) * N + c] * g.igmm1;
  }
  if (MHD) {
    double b[3] = {0.0, 0.0, 0.0};
    b[0] = 0.5 * (g.bf[0][fix(g, 0, i + 1, j, k)] + g.bf[0][fix(g, 0, i, j, k)]);
    b[1] = 0.5 * (g.bf[1][fix(g, 1, i, j + 1, k)] + g.bf[1][fix(g, 1, i, j, k)]);
    if (SD == 3) b[2] = 0.5 * (g.bf[2][fix(g, 2, i, j, k + 1)] + g.bf[2][fix(g, 2, i, j, k)]);
    g.bcc[c] = b[0];
    g.bcc[N + c] = b[1];
    if (SD == 3) g.bcc[2 * N + c] = b[2];
    const double vv[3] = {g.P[N + c], g.P[2 * N + c], SD == 3 ? g.P[3 * N + c] : 0.0};
    if (g.ecellf) {   // mixed path: FP32 cell EMFs
      g.ecellf[2 * N + c] = (float)emf_of<SD>(2, vv, b);
      if (SD == 3) {
        g.ecellf[c] = (float)emf_of<SD>(0, vv, b);
        g.ecellf[N + c] = (float)emf_of<SD>(1, vv, b);
      }
    } else {
      g.ecell[2 * N + c] = emf_of<SD>(2, vv, b);
      if (SD == 3) {
        g.ecell[c] = emf_of<SD>(0, vv, b);
        g.ecell[N + c] = emf_of<SD>(1, vv, b);
      }
    }
  }
}

// max |div B| over cells, reduced like the CFL max.
template <int SD>
__global__ void __launch_bounds__(256) k_divb(Grid g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  double s = 0.0;
  if (i < g.n1 && j < g.n2) {
    double d = (g.bf[0][fix(g, 0, i + 1, j, k)] - g.bf[0][fix(g, 0, i, j, k)]) * g.inv_dx[0] +
               (g.bf[1][fix(g, 1, i, j + 1, k)] - g.bf[1][fix(g, 1, i, j, k)]) * g.inv_dx[1];
    if (SD == 3) d = d + (g.bf[2][fix(g, 2, i, j, k + 1)] - g.bf[2][fix(g, 2, i, j, k)]) * g.inv_dx[2];
    s = fabs(d);
  }
  block_max_to(s, g.red);
}

// ------------------------------------------------------------ fused sweep
// One RK stage's reconstruction, Riemann fluxes and flux divergence in one
// pass (3D, the 2nd-order reconstructions, ideal or FASTEOS states; FP64 or
// the mixed-precision policy).  A block of 256 threads owns a 16 x 15
// (x1, x2) tile of columns and walks a chunk of kSweepK planes along x3.
// Per plane the tile plus a 2-cell halo of sweep inputs (WB deviations,
// cell B, FASTEOS gammas: what load_cell / load_cell_f return) is staged in
// LDS, SoA.  The tile's 17 x 15 x1 faces, 16 x 16 x2 faces and 16 x 15 x3
// faces are three rounds of at most one face per thread: the x1 / x2 faces
// from the LDS stencil, the x3 face (the upper face of the thread's column)
// from a register window of the cells above, each cell loaded and
// reconstructed once along x3 as the column-segment sweep does.  The x1 /
// x2 fluxes go through LDS to the two cells beside each face; the x3 flux
// of the face below is carried in registers.  The divergence keeps the
// order of the multi-kernel update (x1 overwrites, x2 then x3 accumulate;
// the mixed path's FP32 fluxes widened to FP64 first), the face arithmetic
// is flux_face's / flux_face_mx's and every face state that of the
// per-direction sweeps, so the result is bit for bit that of the separate
// sweeps and update.  Face fluxes never reach HBM; the face EMFs and
// contact signs (MHD) are stored for CT (a face on a tile edge is solved by
// both tiles beside it and stored twice with the same bits), and the RK
// combination a1 S0 + a2 S + a3 dt R is written where k_update would write
// it (stage 1 also rebuilds state0 as k_update does, or, mixed, takes it
// from U when the device flags say U is current).  k_update<..., SWEPT>
// then runs cons -> prim, the EoS and the CFL after CT.
// Requires n1 % 16 == 0 (checked by the host); a last x2 tile past n2
// masks its missing rows.
#ifndef PHLX_SWEEP_KC
#define PHLX_SWEEP_KC 32
#endif
constexpr int kSweepK = PHLX_SWEEP_KC, kSweepTy = 15;
template <int D>
struct SweepDir {
  static constexpr int value = D;
};
// waves per SIMD the register allocation aims at (FP64; the FP32 sweeps)
#ifndef PHLX_SWEEP_WAVES
#define PHLX_SWEEP_WAVES 2
#endif
#ifndef PHLX_SWEEP_WAVES_MX
#define PHLX_SWEEP_WAVES_MX 2
#endif

template <int NAS, bool MHD, int EK, bool MIX>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(MIX ? PHLX_SWEEP_WAVES_MX : PHLX_SWEEP_WAVES)))
    k_sweep(Grid g, int first, int fin, double a1, double a2, double c3) {
  constexpr int SD = 3, NV = SD + 2 + NAS, IE = SD + 1, IA = SD + 2;
  constexpr bool GAM = EK == EK_FASTEOS;
  static_assert(EK != EK_HELM_FACE, "per-face Newton states are not staged");
  using Re = typename std::conditional<MIX, float, double>::type;
  constexpr int TX = kSweepT, TY = kSweepTy;     // cells per tile
  constexpr int W = TX + 4, WW = W * (TY + 4);   // LDS row, cells per variable
  // LDS slots per cell: prim (WB deviations), cell B (MHD), gammas (FASTEOS)
  constexpr int OB = NV, OG = NV + (MHD ? 3 : 0), NC = OG + (GAM ? 2 : 0);
  constexpr int NF1 = TY * (TX + 1), NF2 = (TY + 1) * TX;  // x1 / x2 faces per plane
  // two plane tiles where three blocks per CU still fit: plane k+1's
  // interior is then staged while plane k is solved, and no cell of the x3
  // window stays in registers across the in-plane solves
  constexpr int kTileB = NC * WW * (int)sizeof(Re), kFaceB = NV * (NF1 + NF2) * (int)sizeof(Re);
  constexpr bool DB = 2 * kTileB + kFaceB <= 53 * 1024;
  __shared__ Re tile[(DB ? 2 : 1) * NC * WW];
  __shared__ Re f1[NV * NF1];  // x1 faces: [v][row b][face a = 0..16]
  __shared__ Re f2[NV * NF2];  // x2 faces: [v][face b = 0..15][column a]
  if (!g.sc->active) return;
  const int tid = threadIdx.x;
  const int ntx = g.n1 / TX;
  const int i0 = (blockIdx.x % ntx) * TX, j0 = (blockIdx.x / ntx) * TY;
  const int k0 = blockIdx.y * kSweepK, k1 = min(k0 + kSweepK, g.n3);
  const long long N = g.N;
  // thread (a, b): column (i0 + a, j0 + b) for b < TY (the last thread row
  // has no column; it takes the tile's upper x1 faces)
  const bool own = (tid >> 4) < TY;
  auto plane = [&](int kk) { return DB ? tile + (kk & 1) * NC * WW : tile; };

  struct Cell {
    Re q[NV], b[3], ga[2];
  };
  // rows past n2 + 1 (a partial last tile) are read as row n2 + 1: no
  // face inside the grid reaches them, and the ghost map covers two rows
  auto load = [&](int ii, int jj, int kk, Cell &c) {
    jj = min(jj, g.n2 + 1);
    const bool in = ii >= 0 && ii < g.n1 && jj >= 0 && jj < g.n2 && kk >= 0 && kk < g.n3;
    if constexpr (MIX) {
      if (__all(in))
        load_cell_f<SD, NV, MHD, GAM, true>(g, ii, jj, kk, c.q, c.b, c.ga);
      else
        load_cell_f<SD, NV, MHD, GAM, false>(g, ii, jj, kk, c.q, c.b, c.ga);
    } else {
      if (__all(in))
        load_cell<SD, NV, MHD, GAM, true>(g, ii, jj, kk, c.q, c.b, c.ga);
      else
        load_cell<SD, NV, MHD, GAM, false>(g, ii, jj, kk, c.q, c.b, c.ga);
    }
  };
  auto put = [&](Re *t, int cell, const Cell &c) {
#pragma unroll
    for (int v = 0; v < NV; v++) t[v * WW + cell] = c.q[v];
    if (MHD) {
#pragma unroll
      for (int v = 0; v < 3; v++) t[(OB + v) * WW + cell] = c.b[v];
    }
    if (GAM) {
      t[OG * WW + cell] = c.ga[0];
      t[(OG + 1) * WW + cell] = c.ga[1];
    }
  };
  auto states = [&](bool lim, const Re *w, int st, Re &qL, Re &qR) {
    if constexpr (MIX)
      mx::face_states<false>(g.recon, lim, w, st, qL, qR);
    else
      face_states<false>(g.recon, lim, w, st, qL, qR);
  };
  auto edges = [&](bool lim, Re qm, Re q0, Re qp, Re &l, Re &r) {
    if constexpr (MIX)
      cell_edges_mx(g.recon, lim, qm, q0, qp, l, r);
    else
      cell_edges(g.recon, lim, qm, q0, qp, l, r);
  };
  auto solve = [&](auto dirc, int ii, int jj, int kk, Re *qL, Re *qR, Re *bL, Re *bR, Re *gL,
                   Re *gR, Re *Fo) {
    constexpr int DIR = decltype(dirc)::value;
    if constexpr (MIX)
      flux_face_mx<DIR, SD, NAS, MHD, GAM, true>(g, ii, jj, kk, qL, qR, bL, bR, gL, gR, Fo);
    else
      flux_face<DIR, SD, NAS, MHD, EK, false, true>(g, ii, jj, kk, qL, qR, bL, bR, gL, gR, 0.0,
                                                    false, Fo);
  };
  // the in-plane face (DIR 0: lower x1 face of tile cell (fa, fb); DIR 1:
  // lower x2 face) from the LDS stencil, its fluxes into the LDS face array
  auto plane_face = [&](auto dirc, const Re *t, int fa, int fb, int kk) {
    constexpr int DIR = decltype(dirc)::value;
    const int base = DIR == 0 ? (fb + 2) * W + fa : fb * W + fa + 2;
    constexpr int st = DIR == 0 ? 1 : W;
    Re qL[NV], qR[NV], bL[3] = {0, 0, 0}, bR[3] = {0, 0, 0}, gL[2] = {0, 0}, gR[2] = {0, 0};
    Re Fo[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) states(v > IE, &t[v * WW + base], st, qL[v], qR[v]);
    if (MHD) {
#pragma unroll
      for (int c = 0; c < SD; c++) {
        if (c == DIR) continue;
        states(false, &t[(OB + c) * WW + base], st, bL[c], bR[c]);
      }
    }
    if (GAM) {
#pragma unroll
      for (int c = 0; c < 2; c++) states(false, &t[(OG + c) * WW + base], st, gL[c], gR[c]);
    }
    solve(dirc, i0 + fa, j0 + fb, kk, qL, qR, bL, bR, gL, gR, Fo);
    Re *dst = DIR == 0 ? f1 + fb * (TX + 1) + fa : f2 + fb * TX + fa;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      dst[0] = Fo[v];
      dst += DIR == 0 ? NF1 : NF2;
    }
  };

  // prologue: cells k0-2 .. k0+1 of this column; the right edges of cell k0
  // and the flux of face k0 (lower x3 face of the first plane).  A column
  // past n2 (a last, partial tile) only stages its ghost cells.
  Cell nxt;
  Re rq[NV], rb[3] = {0, 0, 0}, rg[2] = {0, 0}, F3m[NV];
  if (own) {
    const int a = tid & (TX - 1), b = tid >> 4, i = i0 + a, j = j0 + b;
    Cell cA, cB, cC;
    load(i, j, k0 - 2, cA);
    load(i, j, k0 - 1, cB);
    load(i, j, k0, cC);
    load(i, j, k0 + 1, nxt);
    Re l, qL[NV], qR[NV], bL[3] = {0, 0, 0}, bR[3] = {0, 0, 0}, gL[2] = {0, 0}, gR[2] = {0, 0};
#pragma unroll
    for (int v = 0; v < NV; v++) {
      edges(v > IE, cA.q[v], cB.q[v], cC.q[v], l, qL[v]);
      edges(v > IE, cB.q[v], cC.q[v], nxt.q[v], qR[v], rq[v]);
    }
    if (MHD) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        edges(false, cA.b[c], cB.b[c], cC.b[c], l, bL[c]);
        edges(false, cB.b[c], cC.b[c], nxt.b[c], bR[c], rb[c]);
      }
    }
    if (GAM) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        edges(false, cA.ga[c], cB.ga[c], cC.ga[c], l, gL[c]);
        edges(false, cB.ga[c], cC.ga[c], nxt.ga[c], gR[c], rg[c]);
      }
    }
    if (j < g.n2) solve(SweepDir<2>{}, i, j, k0, qL, qR, bL, bR, gL, gR, F3m);
    put(plane(k0), (b + 2) * W + a + 2, cC);
  }

  // the RK buffers of k_update: the FP64 path saves state0 in U0 and carries
  // the stage state in U; the mixed path keeps the step's start state in U
  // and the stage state in U0 until the last stage writes U
  double *const s0_buf = MIX ? g.U : g.U0;
  double *const s_buf = MIX ? g.U0 : g.U;
  double *const out = (MIX && !fin) ? g.U0 : g.U;
  const bool from_u = MIX && first && (g.sc->flags & SC_U_CUR);

#pragma unroll 1
  for (int k = k0; k < k1; k++) {
    // the lane's indices re-materialised per plane: hoisting every
    // plane-invariant address and profile value of the face solves out of
    // the loop costs more registers than the kernel has
    int lt = tid;
    asm volatile("" : "+v"(lt));
    const int la = lt & (TX - 1), lb = lt >> 4, li = i0 + la, lj = j0 + lb;
    const bool col = lb < TY && lj < g.n2;  // a column inside the grid
    Re *const tk = plane(k);
    const int mine = (lb + 2) * W + la + 2;
    // double tile: cell k+1 of this column into plane k+1's tile (whose
    // last reads, of plane k-1, are behind the previous iteration's second
    // barrier); the single tile takes it after this plane's reads
    if (DB && lb < TY) put(plane(k + 1), mine, nxt);
    // the halo of plane k (its interior is staged already): x1 halo cells
    // -2, -1, 16, 17 of rows 0..14, x2 halo rows -2, -1, 15, 16
    if (lt < 128 && (lt >= 64 || lt < 4 * TY)) {
      int ha, hb;
      if (lt < 64) {
        const int s = lt & 3;
        ha = s < 2 ? s - 2 : TX - 2 + s;
        hb = lt >> 2;
      } else {
        const int s = (lt - 64) >> 4;
        ha = lt & (TX - 1);
        hb = s < 2 ? s - 2 : TY - 2 + s;
      }
      Cell h;
      load(i0 + ha, j0 + hb, k, h);
      put(tk, (hb + 2) * W + ha + 2, h);
    }
    __syncthreads();
    // in-plane faces, one of each direction per thread: x1 the lower face
    // of cell (la, lb) (the last thread row: the upper face of row la), x2
    // the lower face of row lb (16 faces: row 15's is the tile's upper face)
    {
      const int fa = lb < TY ? la : TX, fb = lb < TY ? lb : la;
      if (fb < TY && j0 + fb < g.n2) plane_face(SweepDir<0>{}, tk, fa, fb, k);
    }
    if (lj <= g.n2) plane_face(SweepDir<1>{}, tk, la, lb, k);
    // cell k+2 of this column, in flight across the barrier
    Cell cp;
    if (lb < TY) load(li, lj, k + 2, cp);
    __syncthreads();
    if (lb < TY) {
      // x3: face k+1 of this column from cells k (LDS), k+1, k+2
      Re qL[NV], qR[NV], bL[3] = {0, 0, 0}, bR[3] = {0, 0, 0}, gL[2] = {0, 0}, gR[2] = {0, 0};
      Re F3p[NV];
      const Re *qm = tk + mine;
      const Re *q1 = plane(k + 1) + mine;
#pragma unroll
      for (int v = 0; v < NV; v++) {
        qL[v] = rq[v];
        edges(v > IE, qm[v * WW], DB ? q1[v * WW] : nxt.q[v], cp.q[v], qR[v], rq[v]);
      }
      if (MHD) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
          bL[c] = rb[c];
          edges(false, qm[(OB + c) * WW], DB ? q1[(OB + c) * WW] : nxt.b[c], cp.b[c], bR[c],
                rb[c]);
        }
      }
      if (GAM) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
          gL[c] = rg[c];
          edges(false, qm[(OG + c) * WW], DB ? q1[(OG + c) * WW] : nxt.ga[c], cp.ga[c], gR[c],
                rg[c]);
        }
      }
      // single tile: plane k's reads are done (this thread's own cell k was
      // the last, just above), so cell k+1 takes its slot
      if (!DB && k + 1 < k1) put(tk, mine, nxt);
      nxt = cp;
      if (col) {
        // the stage's state (state0 and stage state, or state0 rebuilt from
        // the primitives) first: its loads are in flight during the x3 solve
        const long long c = cix(g, li, lj, k);
        double S[NV], S0[NV], R[NV];
        if (from_u) {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            S0[v] = g.U[v * N + c];
            S[v] = S0[v];
          }
        } else if (first) {
          // state0 from the primitives (source/source.F90:7057-7117)
          double p[NV];
#pragma unroll
          for (int v = 0; v < NV; v++) p[v] = g.P[v * N + c];
          const double rho = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
          S0[0] = rho;
          S0[1] = rho * v1;
          S0[2] = rho * v2;
          S0[3] = rho * v3;
          double E = g.eint[c] + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3);
          if (MHD) {
            const double b1 = g.bcc[c], b2 = g.bcc[N + c], b3 = g.bcc[2 * N + c];
            E = E + 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
          }
          if (g.etot) E = E + rho * g.phi_cc[lj];
          S0[IE] = E;
#pragma unroll
          for (int v = IA; v < NV; v++) S0[v] = rho * p[v];
#pragma unroll
          for (int v = 0; v < NV; v++) {
            S[v] = S0[v];
            s0_buf[v * N + c] = S0[v];
          }
        } else {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            S0[v] = s0_buf[v * N + c];
            S[v] = s_buf[v * N + c];
          }
        }
        solve(SweepDir<2>{}, li, lj, k + 1, qL, qR, bL, bR, gL, gR, F3p);
        // divergence (x1 overwrites, x2 and x3 accumulate), sources, RK
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const Re *F1 = f1 + v * NF1 + lb * (TX + 1) + la;
          const Re *F2 = f2 + v * NF2 + lb * TX + la;
          R[v] = ((double)F1[1] - (double)F1[0]) * g.inv_dx[0];
          R[v] = R[v] + ((double)F2[TX] - (double)F2[0]) * g.inv_dx[1];
          R[v] = R[v] + ((double)F3p[v] - (double)F3m[v]) * g.inv_dx[2];
          F3m[v] = F3p[v];
        }
        if (g.grav) {
          const double r0 = S[0];
          const double rho = g.wb ? (r0 - g.eq_rho_cc[lj]) : r0;
          const double gy = g.gravy[lj];
          R[2] = R[2] - rho * gy;
          if (!g.etot) R[IE] = R[IE] - S[2] * gy;
        }
        if (g.edot) R[IE] = R[IE] - g.edot_cc[lj];
        const double a3 = c3 * g.sc->dt;
#pragma unroll
        for (int v = 0; v < NV; v++) out[v * N + c] = a1 * S0[v] + a2 * S[v] + a3 * R[v];
      }
    }
  }
}

// ------------------------------------------------------- shock flattening
// USE_SHOCK_FLATTENING detector (source/source.F90:8073-8131).  k_sf_detect:
// one thread per cell of -2 .. n+1 on each axis marks a cell whose summed
// relative pressure jump sum_d |p(c+e_d) - p(c-e_d)| / p(c) exceeds eps_sf;
// k_sf_dilate: a cell of -1 .. n is flattened when a marked cell lies in
// its 3^sdims neighbourhood (the reference's scatter written as a gather, so
// no flag array has to be cleared first).  Ghost pressures come through the
// boundary map (pressure carries no sign flip); with USE_WB the full
// pressure is the stored deviation plus the mapped row's equilibrium.
template <int SD>
__device__ __forceinline__ double sf_pressure(const Grid &g, int i, int j, int k) {
  int s;
  const int mi = map_axis(g, 0, i, g.n1, s);
  const int mj = map_axis(g, 1, j, g.n2, s);
  const int mk = (SD == 3) ? map_axis(g, 2, k, g.n3, s) : k;
  const double p = g.P[(long long)(SD + 1) * g.N + cix(g, mi, mj, mk)];
  if (!g.wb) return p;
  const double e = g.eq_p_cc[mj];
  return (p - e) + e;
}

// The detector of cell (i, j, k) from the full pressures of its stencil.
template <int SD>
__device__ __forceinline__ unsigned char sf_test(const Grid &g, double pc, double pim,
                                                 double pip, double pjm, double pjp, double pkm,
                                                 double pkp) {
  const double dj = fabs(pjp - pjm), di = fabs(pip - pim);
  double tmp;
  if (SD == 3)
    tmp = (fabs(pkp - pkm) + dj + di) / pc;
  else
    tmp = (dj + di) / pc;
  return (tmp > g.eps_sf) ? 1 : 0;
}

template <int SD>
__global__ void __launch_bounds__(256) k_sf_detect(Grid g) {
  if (!g.sc->active) return;
  int a, b;
  if (!plane_coords(g.n1 + 4, g.n2 + 4, a, b)) return;
  const int i = a - 2, j = b - 2, k = (SD == 3) ? (int)blockIdx.z - 2 : 0;
  const bool inside = i >= 1 && i <= g.n1 - 2 && j >= 1 && j <= g.n2 - 2 &&
                      (SD == 2 || (k >= 1 && k <= g.n3 - 2));
  unsigned char d;
  if (__all(inside)) {
    // every stencil cell of the wave is interior: direct loads
    const double *p = g.P + (long long)(SD + 1) * g.N;
    const long long c = cix(g, i, j, k), sj = g.n1, sk = (long long)g.n1 * g.n2;
    double pc = p[c], pim = p[c - 1], pip = p[c + 1], pjm = p[c - sj], pjp = p[c + sj];
    double pkm = 0.0, pkp = 0.0;
    if (SD == 3) {
      pkm = p[c - sk];
      pkp = p[c + sk];
    }
    if (g.wb) {
      const double e = g.eq_p_cc[j], em = g.eq_p_cc[j - 1], ep = g.eq_p_cc[j + 1];
      pc = (pc - e) + e;
      pim = (pim - e) + e;
      pip = (pip - e) + e;
      pkm = (pkm - e) + e;
      pkp = (pkp - e) + e;
      pjm = (pjm - em) + em;
      pjp = (pjp - ep) + ep;
    }
    d = sf_test<SD>(g, pc, pim, pip, pjm, pjp, pkm, pkp);
  } else {
    d = sf_test<SD>(g, sf_pressure<SD>(g, i, j, k), sf_pressure<SD>(g, i - 1, j, k),
                    sf_pressure<SD>(g, i + 1, j, k), sf_pressure<SD>(g, i, j - 1, k),
                    sf_pressure<SD>(g, i, j + 1, k),
                    SD == 3 ? sf_pressure<SD>(g, i, j, k - 1) : 0.0,
                    SD == 3 ? sf_pressure<SD>(g, i, j, k + 1) : 0.0);
  }
  g.sf_det[sdix(g, i, j, k)] = d;
}

template <int SD>
__global__ void __launch_bounds__(256) k_sf_dilate(Grid g) {
  if (!g.sc->active) return;
  int a, b;
  if (!plane_coords(g.n1 + 2, g.n2 + 2, a, b)) return;
  const int i = a - 1, j = b - 1, k = (SD == 3) ? (int)blockIdx.z - 1 : 0;
  int any = 0;
#pragma unroll
  for (int dk = (SD == 3 ? -1 : 0); dk <= (SD == 3 ? 1 : 0); dk++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
      for (int di = -1; di <= 1; di++) any |= g.sf_det[sdix(g, i + di, j + dj, k + dk)];
  g.sf_flag[sfix(g, i, j, k)] = (unsigned char)(any != 0);
}

// -------------------------------------------------------------- launchers
static inline dim3 grid_for(int e1, int e2, int e3, dim3 blk) {
  return dim3((e1 + blk.x - 1) / blk.x, (e2 + blk.y - 1) / blk.y, e3);
}

// flattened e1 x e2 planes, 256 threads per block (plane_coords)
static inline dim3 grid_plane(int e1, int e2, int e3) {
  const long long nplane = (long long)e1 * e2;
  return dim3((unsigned)((nplane + 255) / 256), 1, e3);
}
static inline dim3 grid_x1_faces(const Grid &g) { return grid_plane(g.n1 + 1, g.n2, g.n3); }
// k_flux_x1c: waves of 63 faces over the n2 (n1 + 2) cells of a plane, 4 per
// block, one plane per blockIdx.y
static inline dim3 grid_x1_cells(const Grid &g) {
  const long long waves = ((long long)(g.n1 + 2) * g.n2 + 62) / 63;
  return dim3((unsigned)((waves + 3) / 4), (unsigned)g.n3, 1);
}
// flat grid of the blocks of grid_for, for kernels that place their blocks
// with xcd_cell_block
static inline dim3 grid_flat(int e1, int e2, int e3, dim3 blk) {
  const dim3 gr = grid_for(e1, e2, e3, blk);
  return dim3(gr.x * gr.y * gr.z, 1, 1);
}

template <int SD, int NAS, bool MHD, int EK, bool MIX, bool R5>
static void launch_flux(const Grid &g, int dir, hipStream_t st) {
  constexpr bool HELM = EK != EK_IDEAL;
  const dim3 blk(64, 4, 1);
  if constexpr (MIX) {
    static_assert(EK != EK_HELM_FACE, "mixed precision needs FASTEOS");
    if (dir == 0 && !R5 && PHLX_FLUX_X1_CELLS) {
      k_flux_x1c<SD, NAS, MHD, EK, true><<<grid_x1_cells(g), 256, 0, st>>>(g);
    } else if (dir == 0) {
      k_flux_mx<0, SD, NAS, MHD, HELM, R5><<<grid_x1_faces(g), blk, 0, st>>>(g);
    } else if (!R5 && PHLX_FLUX_SEG_MX) {
      const int n = dir == 1 ? g.n2 : g.n3, no = dir == 1 ? g.n3 : g.n2;
      const dim3 grd((g.n1 + 63) / 64, (no + 3) / 4, (n + 1 + kSegFacesMx - 1) / kSegFacesMx);
      if (dir == 1)
        k_flux_seg_mx<1, SD, NAS, MHD, HELM><<<grd, blk, 0, st>>>(g);
      else if (SD == 3)
        k_flux_seg_mx<2, SD, NAS, MHD, HELM><<<grd, blk, 0, st>>>(g);
    } else if (dir == 1) {
      k_flux_mx<1, SD, NAS, MHD, HELM, R5><<<grid_for(g.n1, g.n2 + 1, g.n3, blk), blk, 0, st>>>(g);
    } else if (SD == 3) {
      k_flux_mx<2, SD, NAS, MHD, HELM, R5><<<grid_for(g.n1, g.n2, g.n3 + 1, blk), blk, 0, st>>>(g);
    }
  } else {
    if constexpr (!R5 && EK != EK_HELM_FACE) {
      if (dir == 0 && PHLX_FLUX_X1_CELLS) {
        k_flux_x1c<SD, NAS, MHD, EK, false><<<grid_x1_cells(g), 256, 0, st>>>(g);
        return;
      }
    }
    if (dir == 0) {
      k_flux<0, SD, NAS, MHD, EK, R5><<<grid_x1_faces(g), blk, 0, st>>>(g);
    } else if (!R5 && PHLX_FLUX_SEG) {
      // column segments: (x1 cells, transverse cells, segments of kSegFaces)
      const int n = dir == 1 ? g.n2 : g.n3, no = dir == 1 ? g.n3 : g.n2;
      const dim3 grd((g.n1 + 63) / 64, (no + 3) / 4, (n + 1 + kSegFaces - 1) / kSegFaces);
      if (dir == 1)
        k_flux_seg<1, SD, NAS, MHD, EK><<<grd, blk, 0, st>>>(g);
      else if (SD == 3)
        k_flux_seg<2, SD, NAS, MHD, EK><<<grd, blk, 0, st>>>(g);
    } else if (dir == 1) {
      k_flux<1, SD, NAS, MHD, EK, R5><<<grid_for(g.n1, g.n2 + 1, g.n3, blk), blk, 0, st>>>(g);
    } else if (SD == 3) {
      k_flux<2, SD, NAS, MHD, EK, R5><<<grid_for(g.n1, g.n2, g.n3 + 1, blk), blk, 0, st>>>(g);
    }
  }
}

// USE_SHOCK_FLATTENING sweeps (FP64): as launch_flux with the flag-reading
// variants (per-face x1 and LIM5TH sweeps, column segments otherwise).
template <int SD, int NAS, bool MHD, int EK, bool R5>
static void launch_flux_sf(const Grid &g, int dir, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (dir == 0) {
    k_flux<0, SD, NAS, MHD, EK, R5, true><<<grid_x1_faces(g), blk, 0, st>>>(g);
  } else if (!R5 && PHLX_FLUX_SEG) {
    const int n = dir == 1 ? g.n2 : g.n3, no = dir == 1 ? g.n3 : g.n2;
    const dim3 grd((g.n1 + 63) / 64, (no + 3) / 4, (n + 1 + kSegFaces - 1) / kSegFaces);
    if (dir == 1)
      k_flux_seg<1, SD, NAS, MHD, EK, true><<<grd, blk, 0, st>>>(g);
    else if (SD == 3)
      k_flux_seg<2, SD, NAS, MHD, EK, true><<<grd, blk, 0, st>>>(g);
  } else if (dir == 1)
    k_flux<1, SD, NAS, MHD, EK, R5, true><<<grid_for(g.n1, g.n2 + 1, g.n3, blk), blk, 0, st>>>(g);
  else if (SD == 3)
    k_flux<2, SD, NAS, MHD, EK, R5, true><<<grid_for(g.n1, g.n2, g.n3 + 1, blk), blk, 0, st>>>(g);
}

template <int SD, int NAS, bool MHD, bool HELM, bool MIX, bool R5 = false,
          int EK = HELM ? EK_FASTEOS : EK_IDEAL>
struct Kern {
  static void flux(const Grid &g, int dir, hipStream_t st) {
    if constexpr (!MIX) {
      if (g.sflat) {
        launch_flux_sf<SD, NAS, MHD, EK, R5>(g, dir, st);
        return;
      }
    }
    launch_flux<SD, NAS, MHD, EK, MIX, R5>(g, dir, st);
  }
  static void corner(const Grid &g, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_corner<SD, MIX><<<grid_plane(g.n1 + 1, g.n2 + 1, SD == 3 ? g.n3 + 1 : 1), blk, 0, st>>>(g);
  }
  static void faceb(const Grid &g, int mode, double a1, double a2, double c3, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_faceb<SD, MIX><<<grid_plane(g.n1 + 1, g.n2 + 1, SD == 3 ? g.n3 + 1 : 1), blk, 0, st>>>(
        g, mode, a1, a2, c3);
  }
  static void update(const Grid &g, int first, int last, double a1, double a2, double c3,
                     int fin, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_update<SD, NAS, MHD, HELM, MIX><<<grid_flat(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(
        g, first, last, a1, a2, c3, fin);
  }
  static void update_guess(const Grid &g, int first, int last, double a1, double a2, double c3,
                           int fin, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_update<SD, NAS, MHD, HELM, MIX, true><<<grid_flat(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(
        g, first, last, a1, a2, c3, fin);
  }
  // the fused sweep and the update that finishes its stage after CT
  // (3D FP64 with the 2nd-order reconstructions and ideal / FASTEOS states)
  static constexpr bool kSwept = SD == 3 && !R5 && EK != EK_HELM_FACE;
  static void sweep(const Grid &g, int first, int fin, double a1, double a2, double c3,
                    hipStream_t st) {
    if constexpr (kSwept) {
      const dim3 grd((unsigned)((g.n1 / kSweepT) * ((g.n2 + kSweepTy - 1) / kSweepTy)),
                     (unsigned)((g.n3 + kSweepK - 1) / kSweepK), 1);
      k_sweep<NAS, MHD, EK, MIX><<<grd, 256, 0, st>>>(g, first, fin, a1, a2, c3);
    }
  }
  static void update_swept(const Grid &g, int first, int last, double a1, double a2, double c3,
                           int fin, hipStream_t st) {
    if constexpr (kSwept) {
      const dim3 blk(64, 4, 1);
      k_update<SD, NAS, MHD, HELM, MIX, false, true>
          <<<grid_flat(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g, first, last, a1, a2, c3, fin);
    }
  }
  static void update_swept_guess(const Grid &g, int first, int last, double a1, double a2,
                                 double c3, int fin, hipStream_t st) {
    if constexpr (kSwept && HELM) {
      const dim3 blk(64, 4, 1);
      k_update<SD, NAS, MHD, HELM, MIX, true, true>
          <<<grid_flat(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g, first, last, a1, a2, c3, fin);
    }
  }
  static void cfl(const Grid &g, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_cfl<SD, MHD, HELM><<<grid_for(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g);
  }
  static void init(const Grid &g, int restart, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_init<SD, MHD, HELM><<<grid_for(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g, restart);
  }
  static void divb(const Grid &g, hipStream_t st) {
    const dim3 blk(64, 4, 1);
    k_divb<SD><<<grid_for(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g);
  }
  static void prims(const Grid &g, hipStream_t st) {
    if constexpr (MIX) {
      const dim3 blk(64, 4, 1);
      k_prim_from_u<SD, NAS, MHD, HELM><<<grid_for(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g);
    }
  }
  static void pack(const Grid &g, hipStream_t st) {
    if constexpr (MIX) {
      const dim3 blk(64, 4, 1);
      k_pack_f<SD, NAS, MHD, HELM><<<grid_for(g.n1, g.n2, g.n3, blk), blk, 0, st>>>(g);
    }
  }
};

template <int SD, int NAS, bool MHD, bool HELM, bool MIX, bool R5,
          int EK = HELM ? EK_FASTEOS : EK_IDEAL>
static KernelTable make_table() {
  // the LIM5TH families differ only in their sweeps: their tables carry the
  // flux launcher alone and take the rest from the 2nd-order family's unit
  // (so its update, CT and EoS kernels are built once)
  using K = Kern<SD, NAS, MHD, HELM, MIX, R5, EK>;
  KernelTable t{};
  t.flux = &K::flux;
  if constexpr (R5) return t;
  t.corner = &K::corner;
  t.faceb = &K::faceb;
  t.update = &K::update;
  t.update_guess = HELM ? &K::update_guess : nullptr;
  t.cfl = &K::cfl;
  t.init = &K::init;
  t.divb = &K::divb;
  t.pack = MIX ? &K::pack : nullptr;
  t.prims = MIX ? &K::prims : nullptr;
  if constexpr (K::kSwept) {
    t.sweep = &K::sweep;
    t.update_swept = &K::update_swept;
    t.update_swept_guess = HELM ? &K::update_swept_guess : nullptr;
  }
  return t;
}

template <bool MIX, bool R5>
static bool select_prec(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  if (helm) {
    // Helmholtz runs carry Ye and 1/abar as the two advected scalars; PIG
    // runs may carry them (ignored) or none
    if (nas != 2 && nas != 0) return false;
    if (!fasteos) {
      // per-face T(rho, P) Newton in the Riemann solvers: FP64 only
      if (MIX) return false;
      constexpr int EK = EK_HELM_FACE;
      if (nas == 2) {
        if (sd == 2 && !mhd) t = make_table<2, 2, false, true, false, R5, EK>();
        else if (sd == 3 && !mhd) t = make_table<3, 2, false, true, false, R5, EK>();
        else if (sd == 3 && mhd) t = make_table<3, 2, true, true, false, R5, EK>();
        else return false;
      } else {
        if (sd == 2 && !mhd) t = make_table<2, 0, false, true, false, R5, EK>();
        else if (sd == 3 && !mhd) t = make_table<3, 0, false, true, false, R5, EK>();
        else if (sd == 3 && mhd) t = make_table<3, 0, true, true, false, R5, EK>();
        else return false;
      }
      return true;
    }
    if (nas == 2) {
      if (sd == 2 && !mhd) t = make_table<2, 2, false, true, MIX, R5>();
      else if (sd == 3 && !mhd) t = make_table<3, 2, false, true, MIX, R5>();
      else if (sd == 3 && mhd) t = make_table<3, 2, true, true, MIX, R5>();
      else return false;
    } else {
      if (sd == 2 && !mhd) t = make_table<2, 0, false, true, MIX, R5>();
      else if (sd == 3 && !mhd) t = make_table<3, 0, false, true, MIX, R5>();
      else if (sd == 3 && mhd) t = make_table<3, 0, true, true, MIX, R5>();
      else return false;
    }
    return true;
  }
  if (sd == 2 && nas == 0 && !mhd) t = make_table<2, 0, false, false, MIX, R5>();
  else if (sd == 2 && nas == 0 && mhd) t = make_table<2, 0, true, false, MIX, R5>();
  else if (sd == 3 && nas == 0 && !mhd) t = make_table<3, 0, false, false, MIX, R5>();
  else if (sd == 3 && nas == 0 && mhd) t = make_table<3, 0, true, false, MIX, R5>();
  else if (sd == 2 && nas == 2 && !mhd) t = make_table<2, 2, false, false, MIX, R5>();
  else if (sd == 3 && nas == 2 && !mhd) t = make_table<3, 2, false, false, MIX, R5>();
  else if (sd == 3 && nas == 2 && mhd) t = make_table<3, 2, true, false, MIX, R5>();
  else return false;
  return true;
}

}  // namespace phlx
