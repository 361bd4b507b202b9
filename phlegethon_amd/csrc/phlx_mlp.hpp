// phlx_mlp.hpp -- the T(rho, e) surrogate: a 4 -> 64 -> 64 -> 1 tanh MLP
// that supplies the initial guess of the Helmholtz Newton inversion
// (helm_rhoe_given, source/source.F90:24982-25019; the reference starts from
// the carried temperature and has no surrogate).
//
// Inputs (log10 rho, log10 e, abar, zbar) are normalised affinely to about
// [-1, 1]; the output is normalised log10 T.  Two uses: a cold guess
// (log10 T = network output) for callers without a carried temperature, and
// inside the hydro update a corrected warm start,
//   log10 T_g = log10 T_prev + s (y(state now) - y(state at the last solve)),
// where the network's smooth approximation error cancels in the difference
// while the change of T with the new internal energy is kept.  Inference
// runs one wavefront per 64 zones, the layers transposed (hidden units as
// rows, zones as columns) so that each layer's MFMA output is already in the
// register layout of the next layer's B operand: layer 1 (4 -> 64) on
// v_mfma_f32_16x16x16_bf16 with the FP32 inputs split into three BF16 parts
// (hi + mid + lo holds all 24 significand bits, so the products with the
// BF16 weights are those of the FP32 inputs), layer 2 (64 x 64) on gfx950's
// v_mfma_f32_16x16x32_bf16 with the BF16-rounded tanh(h1), layer 3 (64 -> 1)
// as an FP32 dot product finished by a cross-lane butterfly; FP32
// accumulation, FP32 biases.  The exact-table Newton then runs from
// the guess to the reference's tolerance, so the surrogate changes the
// iteration count, never the converged state beyond that tolerance; a guess
// whose inputs leave the training box, or whose first Newton residual
// |e(T_g) - e| / e exceeds a threshold, is replaced by the bracketed guess
// (bisection in log T over the table range).
#pragma once

#include <hip/hip_runtime.h>

#include "phlx_helm.hpp"

namespace phlx {

constexpr int kMlpIn = 4, kMlpH = 64;
// flat parameter vector: w1 [4][64], b1 [64], w2 [64][64], b2 [64], w3 [64], b3
constexpr int kMlpOffW1 = 0, kMlpOffB1 = 256, kMlpOffW2 = 320, kMlpOffB2 = 4416,
              kMlpOffW3 = 4480, kMlpOffB3 = 4544, kMlpNParam = 4545;
// the inference's BF16 A operands in one buffer: w2a, then w1a
constexpr int kMlpW2aShorts = 4096, kMlpW1aShorts = 1024;
constexpr int kBracketSteps = 16;  // bisection steps of the fallback guess
// The inference forms every hidden pre-activation already multiplied by
// 2 / ln 2, the argument of tanh's exp2 (inputs and biases pre-scaled).
constexpr float kMlpExp2Scale = 2.8853900817779268f;

// T(rho, e) starts: the carried T (the reference), the surrogate's
// corrected warm start, the bracketed bisection, and the energy predictor
// T + (e - e_prev) / c_v from the zone's previous solve (its target e_prev
// and slope c_v stored per zone).
enum { GUESS_WARM = 0, GUESS_MLP = 1, GUESS_BRACKET = 2, GUESS_PRED = 3 };

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_bf2 __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_bf8 __attribute__((ext_vector_type(8)));
typedef float mlp_f2 __attribute__((ext_vector_type(2)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

struct MlpDev {
  const float *theta;   // FP32 master parameters (biases are read from here)
  const float *w1q;     // BF16-rounded w1 as FP32, [4][64]
  const float *w3q;     // BF16-rounded w3 as FP32, [64]
  const float *bs;      // b1 [64] then b2 [64], times kMlpExp2Scale
  // MFMA A operands of the transposed layers (lane l: row l % 16, k-slots
  // of lane group q = l / 16):
  //   w1a [hb 4][lane 64] x 4 BF16: w1[e][16 hb + l % 16] for q < 3, 0 for q = 3
  //       (slot group q multiplies the zone's hi / mid / lo input part)
  //   w2a [nb 4][kb 2][lane 64] x 8 BF16: w2[H][16 nb + l % 16] with
  //       H = 32 kb + 4 q + e (e < 4), 32 kb + 16 + 4 q + e - 4 (e >= 4)
  const mlp_s4 *w1a;
  const mlp_bf8 *w2a;
  float in_off[4], in_scale[4];  // x = (raw - off) * scale
  float out_off, out_scale;      // log10 T = y * out_scale + out_off
  float resid_tol;               // fallback when the first residual exceeds this
  float box;                     // accepted |x| (training box + margin)
};

// LDS per wavefront for the inference: the 64 zones' inputs as three BF16
// parts (hi, mid, lo), two dwords each.
constexpr int kMlpLdsWaveDw = 64 * 3 * 2;

__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_round(float f) {
  return __uint_as_float(((unsigned int)bf16_bits(f)) << 16);
}
// tanh(z) = 1 - 2 / (exp(2z) + 1) with the hardware exp2 and reciprocal
// (v_exp_f32, v_rcp_f32: ~1 ulp each), exact limits; the error (~2e-7) is
// far below the BF16 resolution of the activations it feeds.
__device__ __forceinline__ float fast_tanh(float z) {
  z = fminf(fmaxf(z, -15.0f), 15.0f);
  const float e = __builtin_amdgcn_exp2f(z * kMlpExp2Scale);  // 2 / ln 2
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
// The same from u = z * 2 / ln 2.  No clamp: exp2 overflows to +inf for
// large u and the reciprocal of +inf is 0 (tanh = 1), and underflows to 0
// for very negative u (tanh = -1), so the limits are exact as they are.
// 2 r is exact, so the fused 1 - 2 r rounds as the separate operations do.
__device__ __forceinline__ float tanh_exp2(float u) {
  return fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(u) + 1.0f), 1.0f);
}

// log10 of a positive value in FP32 (hardware v_log_f32; the network's
// inputs are FP32); 0 stays finite through the clamp to the FP32 minimum.
__device__ __forceinline__ float mlp_log10f(double v) {
  const float f = fmaxf((float)v, 1.17549435e-38f);
  return __builtin_amdgcn_logf(f) * 0.30102999566398120f;
}

__device__ __forceinline__ void mlp_inputs(const MlpDev &m, double rho, double e, double abar,
                                           double zbar, float x[4], bool &inbox) {
  const float raw[4] = {mlp_log10f(rho), mlp_log10f(fabs(e)), (float)abar, (float)zbar};
  inbox = e > 0.0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    x[i] = (raw[i] - m.in_off[i]) * m.in_scale[i];
    inbox = inbox && fabsf(x[i]) <= m.box;
  }
  if (!inbox) {
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = 0.0f;  // keep the wave's arithmetic finite
  }
}

// The LDS exchange of mlp_forward_wave is within one wavefront (each wave
// owns its kMlpLdsWaveDw dwords), so a wave-scope fence orders it: every
// lane's LDS writes complete before any lane's later reads.  No block
// barrier, so the waves of a block run their inference independently.
__device__ __forceinline__ void mlp_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (a, b) rounded to BF16 (round to nearest even, v_cvt_pk_bf16_f32), a in
// the low half
__device__ __forceinline__ unsigned int mlp_pk_bf16(float a, float b) {
  const mlp_bf2 h = __builtin_convertvector((mlp_f2){a, b}, mlp_bf2);
  return __builtin_bit_cast(unsigned int, h);
}
__device__ __forceinline__ float mlp_bf16_lo(unsigned int u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float mlp_bf16_hi(unsigned int u) {
  return __uint_as_float(u & 0xFFFF0000u);
}

// Normalised network output for this lane's zone.  Every lane of the
// wavefront must call it (MFMA and LDS exchange are wave-wide); lds points
// at this wavefront's kMlpLdsWaveDw dwords, which no other wave touches.
// A layer's MFMA output D (lane l: rows 4q + r, r = 0..3, column l % 16;
// rows = hidden units, columns = zones 16 zb + l % 16) is, after tanh and
// the BF16 rounding, the next layer's B operand in place.
__device__ __forceinline__ float mlp_forward_wave(const MlpDev &m, const float x[4],
                                                  unsigned int *lds) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, col = lane & 15;
  // this zone's inputs, times 2 / ln 2, as hi + mid + lo BF16 parts (each
  // residual is exact)
  {
    unsigned int part[3][2];
#pragma unroll
    for (int pr = 0; pr < 2; pr++) {
      float a = x[2 * pr] * kMlpExp2Scale, b = x[2 * pr + 1] * kMlpExp2Scale;
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const unsigned int h = mlp_pk_bf16(a, b);
        part[t][pr] = h;
        a = a - mlp_bf16_lo(h);
        b = b - mlp_bf16_hi(h);
      }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
      lds[(lane * 3 + t) * 2] = part[t][0];
      lds[(lane * 3 + t) * 2 + 1] = part[t][1];
    }
  }
  mlp_wave_lds_sync();
  // layer-1 B operands: zone 16 zb + col; lane group q < 3 holds part q of
  // its four inputs, group 3 zeros
  mlp_s4 xb[4];
#pragma unroll
  for (int zb = 0; zb < 4; zb++) {
    const unsigned int *src = lds + ((16 * zb + col) * 3 + (q < 3 ? q : 0)) * 2;
    const unsigned int d0 = src[0], d1 = src[1];
    const unsigned long long v =
        q < 3 ? ((unsigned long long)d0 | ((unsigned long long)d1 << 32)) : 0ull;
    xb[zb] = __builtin_bit_cast(mlp_s4, v);
  }
  mlp_wave_lds_sync();
  const float *b1 = m.bs;
  const float *b2 = m.bs + kMlpH;
  // The zone blocks rotate through xb[0] and s[3] (register moves), so the
  // rolled loop indexes no register array by its counter (which would
  // select among the elements lane by lane on every pass).
  float s[4];
#pragma unroll 1
  for (int zb = 0; zb < 4; zb++) {
    // layer 1: h1^T (hidden 16 hb + 4q + r, zone) = w1^T x^T + b1 (times
    // 2 / ln 2), then tanh and BF16 into the k-slots of layer 2 (block
    // kb = hb / 2: slots 0..3 from hb even, 4..7 from hb odd)
    mlp_bf8 h1[2];
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
      const float *bb = b1 + 16 * hb + 4 * q;
      mlp_f4 d = (mlp_f4){bb[0], bb[1], bb[2], bb[3]};
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(m.w1a[hb * 64 + lane], xb[0], d, 0, 0, 0);
      const mlp_bf2 v0 =
          __builtin_bit_cast(mlp_bf2, mlp_pk_bf16(tanh_exp2(d[0]), tanh_exp2(d[1])));
      const mlp_bf2 v1 =
          __builtin_bit_cast(mlp_bf2, mlp_pk_bf16(tanh_exp2(d[2]), tanh_exp2(d[3])));
      const int o = (hb & 1) * 4;
      h1[hb >> 1][o] = v0[0];
      h1[hb >> 1][o + 1] = v0[1];
      h1[hb >> 1][o + 2] = v1[0];
      h1[hb >> 1][o + 3] = v1[1];
    }
    // layer 2, and layer 3 over this lane's 16 output rows
    float acc = 0.0f;
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      mlp_f4 d = (mlp_f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; kb++)
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(m.w2a[(nb * 2 + kb) * 64 + lane], h1[kb], d,
                                                    0, 0, 0);
      const int r0 = 16 * nb + 4 * q;
#pragma unroll
      for (int r = 0; r < 4; r++)
        acc = fmaf(tanh_exp2(fmaf(d[r], kMlpExp2Scale, b2[r0 + r])), m.w3q[r0 + r], acc);
    }
    xb[0] = xb[1];
    xb[1] = xb[2];
    xb[2] = xb[3];
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    s[3] = acc;  // after the last pass s[b] holds zone block b
  }
  // y(zone 16 zb + col) = b3 + the sum of s[zb] over the four lane groups;
  // lane l keeps zb = q: halves exchanged across groups q ^ 2, then q ^ 1
  const bool up = (q & 2) != 0, odd = (q & 1) != 0;
  const float keep0 = up ? s[2] : s[0], keep1 = up ? s[3] : s[1];
  const float send0 = up ? s[0] : s[2], send1 = up ? s[1] : s[3];
  const float t0 = keep0 + __shfl_xor(send0, 32);
  const float t1 = keep1 + __shfl_xor(send1, 32);
  const float u = (odd ? t1 : t0) + __shfl_xor(odd ? t0 : t1, 16);
  return m.theta[kMlpOffB3] + u;
}

// Fallback: bisection of e(T) - e0 in log10 T over [Tl, Th].
__device__ __forceinline__ double helm_bracket_guess(const HelmDev &h, double rho, double E0,
                                                     double abar, double zbar) {
  double lo = log10(h.tl), hi = log10(h.th);
#pragma unroll 1
  for (int it = 0; it < kBracketSteps; it++) {
    const double mid = 0.5 * (lo + hi);
    double e, dedt;
    helm_given2(h, rho, exp10(mid), abar, zbar, e, dedt, false);
    if (e > E0) hi = mid; else lo = mid;
  }
  return exp10(0.5 * (lo + hi));
}

// The reference's Newton loop for T(rho, e) (helm_rhoe_given) entered with
// the first evaluation (e1, dedt1 = E and dE/dT at the incoming T) already
// done: identical arithmetic to helm_rhoe_given from the same T, without
// repeating that evaluation.
__device__ __forceinline__ int helm_rhoe_given_from(const HelmDev &h, double rho, double E0,
                                                    double abar, double zbar, double &T,
                                                    double &P, double &sound, bool ret,
                                                    double e1, double dedt1) {
  const double Told = T;
  int iter;
  sound = 0.0;
  for (iter = 1; iter <= kHelmMaxIter; iter++) {
    double e = e1, dEdT = dedt1;
    if (iter > 1) helm_given2(h, rho, T, abar, zbar, e, dEdT, true);
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

// Mixed-precision T(rho, e), linearised (opt-in, phlx_set_eos_lin_tol > 0;
// an algorithmic shortcut, not the reference's convergence test): ONE FP64
// evaluation (helm_rhoT_given_full plus dP/dT) at the start T.  When its energy residual |e(T) - e0| / e0 is
// below lin_tol, T and P take that Newton step linearly, dT = -(e - e0) / c_v,
// T += dT, P += dP/dT dT, which leaves an error of second order in the
// residual (~1e-10 relative for a 1e-5 residual); the sound speed stays that
// of the start (relative error ~dT/T, it feeds the FP32 wave speeds only).
// Otherwise the reference Newton (helm_rhoe_given) runs from the stepped T.
// Returns the EoS evaluations before the final one; fallback = 1 when the
// reference loop ran.
__device__ __forceinline__ int helm_rhoe_lin(const HelmDev &h, double lin_tol, double rho,
                                             double E0, double abar, double zbar, double &T,
                                             double &P, double &sound, int &fallback,
                                             double &cv_out) {
  double E, cv, dPdT;
  helm_full<true>(h, rho, T, abar, zbar, P, E, sound, cv, &dPdT);
  cv_out = cv;
  const double res = E - E0;
  const double dT = hdiv<true>(-res, cv);
  const double T1 = T + dT;
  const bool inside = (T1 >= h.tl) && (T1 <= h.th);
  if (fabs(hdiv<true>(res, E0)) < lin_tol && inside) {
    T = T1;
    P = P + dPdT * dT;
    fallback = 0;
    return 1;
  }
  fallback = 1;
  if (inside) T = T1;
  return 1 + helm_rhoe_given(h, rho, E0, abar, zbar, T, P, sound, true);
}

// Mixed-precision T(rho, e) converged to the reference's tolerance: the
// Newton loop of helm_rhoe_given (T -= (e(T) - e0) / c_v; the loop ends once
// |e(T) - e0| / e0 < helm_tol, the step of that iteration applied; a T
// outside the table resets to the start and ends it), every iterate an FP64
// evaluation of the table interpolant from ONE contraction of the cell over
// rho (HelmMx; refined reciprocals, FP32 cell lookup).  The reference closes
// with a helm_full at the final T; here P takes the final step linearly,
// P(T_k) + dP/dT (T - T_k): that step is below the tolerance, so the
// linearisation error (~(dT / T)^2 ~ 1e-22 relative) is below FP64
// resolution, and the sound speed is the last iterate's (relative
// difference ~dT / T; it feeds the FP32 wave speeds and the CFL).  After a
// reset the closing evaluation is made at the start T.  Returns the Newton
// iterations (as helm_rhoe_given); cv_out receives the last iterate's c_v.
__device__ __forceinline__ int helm_rhoe_mixed(const HelmDev &h, double rho, double E0,
                                               double abar, double zbar, double &T, double &P,
                                               double &sound, double &cv_out) {
  const double Told = T;
  HelmMx m;
  helm_mx_setup(h, rho, T, abar, zbar, m);
  int it = 0;
#pragma unroll 1
  for (;;) {
    double E, cv;
    HelmMxAt at;
    helm_mx_energy(h, m, rho, T, at, E, cv);
    it++;
    const double res = E - E0;
    double Tn = T - hdiv<true>(res, cv);
    double err = fabs(hdiv<true>(res, E0));
    bool reset = false;
    if ((Tn < h.tl) || (Tn > h.th)) {
      Tn = Told;
      err = 1e-15;
      reset = true;
    }
    cv_out = cv;
    if (err < kHelmTol || it >= kHelmMaxIter) {
      // P, dP/dT and the sound speed only at the last iterate
      if (reset) {
        double e2, cv2, d2;
        HelmMxAt a2;
        helm_mx_energy(h, m, rho, Tn, a2, e2, cv2);
        helm_mx_pressure(m, rho, Tn, a2, e2, cv2, P, d2, sound);
      } else {
        double Pk, dPdT;
        helm_mx_pressure(m, rho, T, at, E, cv, Pk, dPdT, sound);
        P = Pk + dPdT * (Tn - T);
      }
      T = Tn;
      return it;
    }
    T = Tn;
  }
}

// The energy predictor's start for a zone whose previous solve had target
// e_prev, slope c_v and result T: T + (e0 - e_prev) / c_v (the previous
// residual is below the tolerance, so the error is second order in the
// change of e).  A missing record (NaN) or a predicted change above 10 %
// keeps T.
__device__ __forceinline__ double helm_pred_start(double T, double e0, double e_prev,
                                                  float cvf) {
  const double cv = (double)cvf;
  if (!(cv > 0.0) || !(e_prev == e_prev)) return T;
  const double tg = T + hdiv<true>(e0 - e_prev, cv);
  return (fabs(tg - T) <= 0.1 * T) ? tg : T;
}

// T(rho, e) with a chosen start (GUESS_WARM: T as given, the reference;
// GUESS_BRACKET: the bisection guess; GUESS_MLP: tg, the surrogate's guess).
// A surrogate guess is judged by the first Newton evaluation itself: if it
// is outside the table or |e(tg) - e| / e > resid_tol the solve restarts
// from the bracketed guess (fallback = 1).  Returns the Newton iterations of
// the final solve.
__device__ __forceinline__ int helm_rhoe_guess(const HelmDev &h, int mode, double tg,
                                               float resid_tol, double rho, double E0,
                                               double abar, double zbar, double &T, double &P,
                                               double &sound, bool ret, int &fallback,
                                               double *cv_last = nullptr) {
  fallback = 0;
  if (mode == GUESS_PRED) {
    if (tg >= h.tl && tg <= h.th) T = tg;
    return helm_rhoe_given(h, rho, E0, abar, zbar, T, P, sound, ret, cv_last);
  }
  if (mode == GUESS_MLP) {
    // a guess outside the table's T range is clamped into it (in degenerate
    // matter log T is ill-determined and any T there has a small energy
    // residual); a missing guess (tg <= 0 or NaN) falls back
    if (tg > 0.0) {
      tg = fmin(fmax(tg, h.tl * (1.0 + 1e-9)), h.th * (1.0 - 1e-9));
      double e, dedt;
      helm_given2(h, rho, tg, abar, zbar, e, dedt, true);
      if (fabs((e - E0) / E0) <= (double)resid_tol) {
        T = tg;
        return helm_rhoe_given_from(h, rho, E0, abar, zbar, T, P, sound, ret, e, dedt);
      }
    }
    fallback = 1;
    T = helm_bracket_guess(h, rho, E0, abar, zbar);
  } else if (mode == GUESS_BRACKET) {
    T = helm_bracket_guess(h, rho, E0, abar, zbar);
  }
  return helm_rhoe_given(h, rho, E0, abar, zbar, T, P, sound, ret);
}

}  // namespace phlx
