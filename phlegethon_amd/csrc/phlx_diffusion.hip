// phlx_diffusion.hip -- THERMAL_DIFFUSION_STS: RKL2 super-time-stepping of
// the radiative diffusion of internal energy, Strang-split around the hydro
// step (source/source.F90:6672-6688, thermal_diffusion_step :19949-20269).
//
// One diffusion step over dt/2 with s stages (s from dt / dt_par on the host):
//   [k_diff_t_from_p]            ideal gas: T of the hydro state
//   k_diff_stage<FIRST>          residual M(T), e_0 = eint, Me_0 = M, stage 1
//   s-1 x (k_diff_temp, k_diff_stage)   eint_to_temp, residual, RKL2 stage
//   k_diff_final                 eint_to_temp, p (Helmholtz: eint, gammas)
// The residual is sum_d (F(+) - F(-)) / dx_d of the face flux
// F = (4/3) a c K_int dT/dx, K = T^3 / (rho kappa), K_int the harmonic
// mean of the two cells; each face flux is recomputed by the two cells that
// share it (same operands, same bits) instead of a face array round trip.
// Every kernel is one thread per cell and HBM bound (7-point stencils of T,
// rho and kappa from cache).
#include "phlx_internal.hpp"

namespace phlx {

constexpr double kRad = 7.565767381646406e-15;  // source/source.F90:614
constexpr double kClight = 2.99792458e10;        // source/source.F90:622

__device__ __forceinline__ long long kapix(const Grid &g, int i, int j, int k) {
  const int k0 = (g.sd == 3) ? 1 : 0;
  return ((long long)(k + k0) * (g.n2 + 2) + (j + 1)) * (g.n1 + 2) + (i + 1);
}

// (T, rho) of the cell beside (i, j, k) along axis a at offset o = -1/+1 (the
// cell itself at o = 0); ghosts per bcs_thermal_diffusion (:20837-21068):
// periodic copies, else rho extrapolated linearly from the wall cells and T
// copied from the wall cell or set to 2 T_wall - T (FIX_TEMPERATURE_AT_*).
template <int SD>
__device__ __forceinline__ void diff_cell(const Grid &g, int i, int j, int k, int a, int o,
                                          double &T, double &rho) {
  int x[3] = {i, j, k};
  const int n[3] = {g.n1, g.n2, g.n3};
  x[a] += o;
  const double *tp = g.tdiff;
  if (x[a] >= 0 && x[a] < n[a]) {
    const long long c = cix(g, x[0], x[1], x[2]);
    T = tp[c];
    rho = g.P[c];
    return;
  }
  const int side = x[a] < 0 ? 0 : 1;
  if (g.bc[2 * a + side] == BC_PERIODIC) {
    x[a] = side ? x[a] - n[a] : x[a] + n[a];
    const long long c = cix(g, x[0], x[1], x[2]);
    T = tp[c];
    rho = g.P[c];
    return;
  }
  x[a] = side ? n[a] - 1 : 0;
  const long long c0 = cix(g, x[0], x[1], x[2]);
  x[a] = side ? n[a] - 2 : 1;
  const long long c1 = cix(g, x[0], x[1], x[2]);
  rho = 2.0 * g.P[c0] - g.P[c1];
  const double tw = g.fix_t[2 * a + side];
  T = (tw == tw) ? 2.0 * tw - tp[c0] : tp[c0];
}

// Radiative flux of the face between cells m and p (:20488-20509).
__device__ __forceinline__ double diff_face(double tm, double rm, double km, double tp,
                                            double rp, double kp, double inv_dl) {
  const double krad = (4.0 / 3.0) * kRad * kClight;
  const double kmn = tm * tm * tm / (rm * km);
  const double kpl = tp * tp * tp / (rp * kp);
  const double kint = 2.0 * (kmn * kpl) / (kmn + kpl);
  const double grad = (tp - tm) * inv_dl;
  return krad * kint * grad;
}

// residuals_thermal_diffusion (:20394-20835), Cartesian, x1 first.
template <int SD>
__device__ __forceinline__ double diff_residual(const Grid &g, int i, int j, int k) {
  const long long c = cix(g, i, j, k);
  const double tc = g.tdiff[c], rc = g.P[c], kc = g.kappa[kapix(g, i, j, k)];
  double m = 0.0;
#pragma unroll
  for (int a = 0; a < SD; a++) {
    double tm, rm, tq, rq;
    diff_cell<SD>(g, i, j, k, a, -1, tm, rm);
    diff_cell<SD>(g, i, j, k, a, +1, tq, rq);
    const double km = g.kappa[kapix(g, i - (a == 0), j - (a == 1), k - (a == 2))];
    const double kq = g.kappa[kapix(g, i + (a == 0), j + (a == 1), k + (a == 2))];
    const double fm = diff_face(tm, rm, km, tc, rc, kc, g.inv_dx[a]);
    const double fp = diff_face(tc, rc, kc, tq, rq, kq, g.inv_dx[a]);
    const double d = (fp - fm) * g.inv_dx[a];
    m = (a == 0) ? d : m + d;
  }
  return m;
}

template <int SD>
__device__ __forceinline__ bool diff_coords(const Grid &g, int &i, int &j, int &k) {
  i = blockIdx.x * blockDim.x + threadIdx.x;
  j = blockIdx.y * blockDim.y + threadIdx.y;
  k = blockIdx.z;
  return i < g.n1 && j < g.n2;
}

// Ideal gas: T of the hydro state, p / (rho R / mu) (:13670-13674).
template <int SD>
__global__ void __launch_bounds__(256) k_diff_t_from_p(Grid g) {
  if (!g.sc->active) return;
  int i, j, k;
  if (!diff_coords<SD>(g, i, j, k)) return;
  const long long c = cix(g, i, j, k);
  g.tdiff[c] = g.P[(long long)(SD + 1) * g.N + c] / (g.P[c] * kRgas * g.inv_mu);
}

// Stage 1 (FIRST): e_0 = e_jm2 = eint, Me_0 = M, eint = e_0 + (dt/2 mu~_1) Me_0.
// Stage j: eint = mu e_jm1 + nu e_jm2 + (1 - mu - nu) e_0 + (dt/2 mu~) M
//                 + (dt/2 gamma~) Me_0, e_jm2 = e_jm1 (:20129-20190).
template <int SD, bool FIRST>
__global__ void __launch_bounds__(256) k_diff_stage(Grid g, double mu, double nu, double c0,
                                                    double hm, double hg) {
  if (!g.sc->active) return;
  int i, j, k;
  if (!diff_coords<SD>(g, i, j, k)) return;
  const long long c = cix(g, i, j, k);
  const double m = diff_residual<SD>(g, i, j, k);
  if (FIRST) {
    const double e0 = g.eint[c];
    g.e0[c] = e0;
    g.me0[c] = m;
    g.ejm2[c] = e0;
    g.eint[c] = e0 + hm * m;
  } else {
    const double ejm1 = g.eint[c];
    g.eint[c] = mu * ejm1 + nu * g.ejm2[c] + c0 * g.e0[c] + hm * m + hg * g.me0[c];
    g.ejm2[c] = ejm1;
  }
}

// eint_to_temp (:20271-20392): the ideal gas in closed form, Helmholtz by
// helm_rhoe_given from the carried T.
template <int SD, bool HELM>
__device__ __forceinline__ double diff_eint_to_temp(const Grid &g, long long c, double rho,
                                                    double eint) {
  if (!HELM) return g.gmm1 * eint / (g.inv_mu * rho * kRgas);
  double ye, abar;
  cell_ye_abar(g, c, ye, abar);
  const double zbar = ye * abar;
  double T = g.tdiff[c], pr, sound;
  helm_rhoe_given(g.helm, rho, eint / rho, abar, zbar, T, pr, sound, false);
  return T;
}

template <int SD, bool HELM>
__global__ void __launch_bounds__(256) k_diff_temp(Grid g) {
  if (!g.sc->active) return;
  int i, j, k;
  if (!diff_coords<SD>(g, i, j, k)) return;
  const long long c = cix(g, i, j, k);
  g.tdiff[c] = diff_eint_to_temp<SD, HELM>(g, c, g.P[c], g.eint[c]);
}

// The step's end (:20193-20267): eint_to_temp, then p from (rho, T); with
// Helmholtz also eint = rho e(T) and the FASTEOS gammas.
template <int SD, bool HELM>
__global__ void __launch_bounds__(256) k_diff_final(Grid g) {
  if (!g.sc->active) return;
  int i, j, k;
  if (!diff_coords<SD>(g, i, j, k)) return;
  constexpr int IP = SD + 1;
  const long long c = cix(g, i, j, k), N = g.N;
  const double rho = g.P[c];
  const double T = diff_eint_to_temp<SD, HELM>(g, c, rho, g.eint[c]);
  g.tdiff[c] = T;
  if (HELM) {
    double ye, abar;
    cell_ye_abar(g, c, ye, abar);
    const double zbar = abar * ye;
    double pr, e, sound, cv;
    helm_full(g.helm, rho, T, abar, zbar, pr, e, sound, cv);
    helm_pin(pr);
    helm_pin(sound);
    const double eint = rho * e;
    g.P[IP * N + c] = pr;
    g.eint[c] = eint;
    if (g.fasteos) {
      g.gam[c] = pr / eint + 1.0;
      g.gam[N + c] = sound * sound * rho / pr;
    }
  } else {
    g.P[IP * N + c] = kRgas * rho * T * g.inv_mu;
  }
}

// compute_parabolic_dt (:19804-19947) without USE_TIMMES_KAPPA: per-block
// partial maxima of 2 D sum 1/dx^2, D = c / (3 kappa rho).
template <int SD>
__global__ void __launch_bounds__(256) k_parabolic(Grid g, double inv_dl2) {
  __shared__ double red[4];
  int i, j, k;
  double v = 0.0;
  if (diff_coords<SD>(g, i, j, k)) {
    const double rho = g.P[cix(g, i, j, k)];
    const double drad = (1.0 / 3.0) * kClight / (g.kappa[kapix(g, i, j, k)] * rho);
    v = 2.0 * drad * inv_dl2;
  }
  // block maximum (max is exact, so the order does not matter)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if (threadIdx.x == 0) red[threadIdx.y] = v;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    g.red[blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = m;
  }
}

static dim3 diff_grid(const Grid &g) {
  return dim3((g.n1 + 63) / 64, (g.n2 + 3) / 4, g.n3);
}

void launch_diff_t_from_p(const Grid &g, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) k_diff_t_from_p<3><<<diff_grid(g), blk, 0, st>>>(g);
  else k_diff_t_from_p<2><<<diff_grid(g), blk, 0, st>>>(g);
}

void launch_diff_stage(const Grid &g, bool first, double mu, double nu, double c0, double hm,
                       double hg, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) {
    if (first) k_diff_stage<3, true><<<diff_grid(g), blk, 0, st>>>(g, mu, nu, c0, hm, hg);
    else k_diff_stage<3, false><<<diff_grid(g), blk, 0, st>>>(g, mu, nu, c0, hm, hg);
  } else {
    if (first) k_diff_stage<2, true><<<diff_grid(g), blk, 0, st>>>(g, mu, nu, c0, hm, hg);
    else k_diff_stage<2, false><<<diff_grid(g), blk, 0, st>>>(g, mu, nu, c0, hm, hg);
  }
}

void launch_diff_temp(const Grid &g, bool helm, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) {
    if (helm) k_diff_temp<3, true><<<diff_grid(g), blk, 0, st>>>(g);
    else k_diff_temp<3, false><<<diff_grid(g), blk, 0, st>>>(g);
  } else {
    if (helm) k_diff_temp<2, true><<<diff_grid(g), blk, 0, st>>>(g);
    else k_diff_temp<2, false><<<diff_grid(g), blk, 0, st>>>(g);
  }
}

void launch_diff_final(const Grid &g, bool helm, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) {
    if (helm) k_diff_final<3, true><<<diff_grid(g), blk, 0, st>>>(g);
    else k_diff_final<3, false><<<diff_grid(g), blk, 0, st>>>(g);
  } else {
    if (helm) k_diff_final<2, true><<<diff_grid(g), blk, 0, st>>>(g);
    else k_diff_final<2, false><<<diff_grid(g), blk, 0, st>>>(g);
  }
}

void launch_parabolic(const Grid &g, double inv_dl2, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) k_parabolic<3><<<diff_grid(g), blk, 0, st>>>(g, inv_dl2);
  else k_parabolic<2><<<diff_grid(g), blk, 0, st>>>(g, inv_dl2);
}

}  // namespace phlx
