// phlx_helm.hip -- batched Helmholtz EoS kernels (one thread per point).
//
// Batched device forms of helm_rhoT_given_full, helm_rhoe_given and
// helm_rhoP_given (source/source.F90:23686-24232, :24943-25019): the
// reference calls them per zone from cons->prim and per face state from the
// Riemann solver; here a caller hands over n points resident in HBM.
#include "phlx_internal.hpp"

namespace phlx {

namespace {

constexpr int kEosBlock = 256;

__global__ void __launch_bounds__(kEosBlock)
k_helm_full(HelmDev h, long long n, const double *__restrict__ rho,
            const double *__restrict__ T, const double *__restrict__ abar,
            const double *__restrict__ zbar, double *__restrict__ P, double *__restrict__ E,
            double *__restrict__ cs, double *__restrict__ cv) {
  const long long i = (long long)blockIdx.x * kEosBlock + threadIdx.x;
  if (i >= n) return;
  double p, e, s, c;
  helm_full(h, rho[i], T[i], abar[i], zbar[i], p, e, s, c);
  helm_pin(p);
  helm_pin(e);
  helm_pin(s);
  helm_pin(c);
  if (P) P[i] = p;
  if (E) E[i] = e;
  if (cs) cs[i] = s;
  if (cv) cv[i] = c;
}

__global__ void __launch_bounds__(kEosBlock)
k_helm_rhoe(HelmDev h, long long n, const double *__restrict__ rho,
            const double *__restrict__ e, const double *__restrict__ abar,
            const double *__restrict__ zbar, double *__restrict__ T, double *__restrict__ P,
            double *__restrict__ cs, int ret, int *__restrict__ iters) {
  const long long i = (long long)blockIdx.x * kEosBlock + threadIdx.x;
  if (i >= n) return;
  double t = T[i], p, s;
  const int it = helm_rhoe_given(h, rho[i], e[i], abar[i], zbar[i], t, p, s, ret != 0);
  helm_pin(p);
  helm_pin(s);
  T[i] = t;
  if (P) P[i] = p;
  if (cs) cs[i] = s;
  if (iters) iters[i] = it;
}

__global__ void __launch_bounds__(kEosBlock)
k_helm_rhoP(HelmDev h, long long n, const double *__restrict__ rho,
            const double *__restrict__ p, const double *__restrict__ abar,
            const double *__restrict__ zbar, double *__restrict__ T, double *__restrict__ E,
            double *__restrict__ cs, int ret, int *__restrict__ iters) {
  const long long i = (long long)blockIdx.x * kEosBlock + threadIdx.x;
  if (i >= n) return;
  double t = T[i], e, s;
  const int it = helm_rhoP_given(h, rho[i], p[i], abar[i], zbar[i], t, e, s, ret != 0);
  helm_pin(e);
  helm_pin(s);
  T[i] = t;
  if (E) E[i] = e;
  if (cs) cs[i] = s;
  if (iters) iters[i] = it;
}

// file column order -> record slot: f, fd, ft, fdd, ftt, fdt, fddt, fdtt,
// fddtt -> 0 3 1 4 2 5 6 7 8; c, cd, ct, cdt -> 9 11 10 12.
__global__ void k_helm_pack(long long np, const double *__restrict__ f9,
                            const double *__restrict__ c4, double *__restrict__ tab) {
  const long long p = (long long)blockIdx.x * kEosBlock + threadIdx.x;
  if (p >= np) return;
  constexpr int FS[9] = {0, 3, 1, 4, 2, 5, 6, 7, 8};
  constexpr int CS[4] = {9, 11, 10, 12};
  double *r = tab + (size_t)kHelmRec * p;
#pragma unroll
  for (int v = 0; v < 9; v++) r[FS[v]] = f9[9 * p + v];
#pragma unroll
  for (int v = 0; v < 4; v++) r[CS[v]] = c4[4 * p + v];
  r[13] = 0.0;
  r[14] = 0.0;
  r[15] = 0.0;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kEosBlock - 1) / kEosBlock); }

}  // namespace

void launch_helm_pack(long long np, const double *f9, const double *c4, double *tab,
                      hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_helm_pack, dim3(blocks_for(np)), dim3(kEosBlock), 0, st, np, f9, c4, tab);
}

void launch_helm_full(const HelmDev &h, long long n, const double *rho, const double *T,
                      const double *abar, const double *zbar, double *P, double *E, double *cs,
                      double *cv, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_helm_full, dim3(blocks_for(n)), dim3(kEosBlock), 0, st, h, n, rho, T,
                     abar, zbar, P, E, cs, cv);
}

void launch_helm_rhoe(const HelmDev &h, long long n, const double *rho, const double *e,
                      const double *abar, const double *zbar, double *T, double *P, double *cs,
                      int ret, int *iters, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_helm_rhoe, dim3(blocks_for(n)), dim3(kEosBlock), 0, st, h, n, rho, e,
                     abar, zbar, T, P, cs, ret, iters);
}

void launch_helm_rhoP(const HelmDev &h, long long n, const double *rho, const double *p,
                      const double *abar, const double *zbar, double *T, double *E, double *cs,
                      int ret, int *iters, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_helm_rhoP, dim3(blocks_for(n)), dim3(kEosBlock), 0, st, h, n, rho, p,
                     abar, zbar, T, E, cs, ret, iters);
}

}  // namespace phlx
