// phlx_mlp.hip -- training and batched inference of the T(rho, e)
// surrogate (phlx_mlp.hpp), and the batched Newton with a selectable guess.
//
// Training data are generated on the GPU by the Helmholtz EoS itself:
// (log10 rho*Ye, log10 T, abar, Ye) drawn uniformly over the table box,
// e = E(rho, T) from the table, and the label log10 T from the FP64 Newton
// solve of T(rho, e) started at the bracketed guess.  The loss weights each
// sample's squared log T error by (d ln E / d ln T)^2 at the label, i.e. it
// measures the guess in the energy residual the Newton iteration works on:
// in degenerate matter E hardly depends on T (d ln T / d ln E reaches 1e4),
// so log T there is ill-determined and irrelevant to the iteration count.  One optimisation step
// is two kernels: k_mlp_grad (128 samples per block: forward with the same
// BF16 rounding points as the inference, backward in FP32, block-partial
// gradients reduced through LDS) and k_mlp_adam (sums the partials, Adam
// update of the FP32 master weights, refreshes the BF16 copies and the MFMA
// B tiles of w2).
#include "phlx_internal.hpp"

namespace phlx {

namespace {

constexpr int kGradBlock = 128;
constexpr int kAdamBlock = 256;
constexpr int kFwdBlock = 256;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  // splitmix64 finaliser (Steele, Lea & Flood 2014)
  z = z + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit01(unsigned long long seed, long long i, int c) {
  const unsigned long long b = mix64(seed * 0x100000001B3ull + (unsigned long long)i * 4ull + c);
  return (double)(b >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256)
k_mlp_data(HelmDev h, long long n, unsigned long long seed, double lrs_lo, double lrs_hi,
           double lt_lo, double lt_hi, double a_lo, double a_hi, double ye_lo, double ye_hi,
           float *__restrict__ raw, float *__restrict__ label, float *__restrict__ wt) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double lrs = lrs_lo + unit01(seed, i, 0) * (lrs_hi - lrs_lo);
  const double lt = lt_lo + unit01(seed, i, 1) * (lt_hi - lt_lo);
  const double abar = a_lo + unit01(seed, i, 2) * (a_hi - a_lo);
  const double ye = ye_lo + unit01(seed, i, 3) * (ye_hi - ye_lo);
  const double rho = exp10(lrs) / ye, zbar = ye * abar;
  double e, dedt;
  helm_given2(h, rho, exp10(lt), abar, zbar, e, dedt, false);
  double T = helm_bracket_guess(h, rho, e, abar, zbar), P, cs;
  helm_rhoe_given(h, rho, e, abar, zbar, T, P, cs, false);
  double el, dedtl;
  helm_given2(h, rho, T, abar, zbar, el, dedtl, true);
  const double slope = dedtl * T / el;  // d ln E / d ln T at the label
  wt[i] = (float)fmin(fmax(slope * slope, 1e-8), 1e4);
  raw[4 * i + 0] = (float)log10(rho);
  raw[4 * i + 1] = (float)log10(e);
  raw[4 * i + 2] = (float)abar;
  raw[4 * i + 3] = (float)zbar;
  label[i] = (float)log10(T);
}

__global__ void __launch_bounds__(256)
k_mlp_normalize(MlpDev m, long long n, const float *__restrict__ raw,
                const float *__restrict__ label, float *__restrict__ xn, float *__restrict__ tn) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 4; c++) xn[4 * i + c] = (raw[4 * i + c] - m.in_off[c]) * m.in_scale[c];
  tn[i] = (label[i] - m.out_off) / m.out_scale;
}

// Per-block partial gradient of the mean squared error over samples
// [n0, n0 + nb) (inv_b = 1 / batch size).  Parameter order as kMlpOff*.
__global__ void __launch_bounds__(kGradBlock)
k_mlp_grad(const float *__restrict__ theta, const float *__restrict__ w1q,
           const float *__restrict__ w3q, const float *__restrict__ xn,
           const float *__restrict__ tn, const float *__restrict__ wt, long long n0,
           long long nb, float inv_b, float *__restrict__ gpart, float *__restrict__ lpart) {
  __shared__ float w2s[kMlpH * kMlpH];
  __shared__ float A[kGradBlock * 65];
  __shared__ float Bm[kGradBlock * 65];
  __shared__ float xs[kGradBlock * 4];
  __shared__ float d3s[kGradBlock];
  const int tid = threadIdx.x;
  for (int e = tid; e < kMlpH * kMlpH; e += kGradBlock) w2s[e] = bf16_round(theta[kMlpOffW2 + e]);
  const long long sl = (long long)blockIdx.x * kGradBlock + tid;
  const bool valid = sl < nb;
  const long long s = n0 + sl;
  float x[4], t = 0.0f, w = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; c++) x[c] = valid ? xn[4 * s + c] : 0.0f;
  if (valid) {
    t = tn[s];
    w = wt ? wt[s] : 1.0f;
  }
  float h1[kMlpH], h2[kMlpH];
#pragma unroll
  for (int k = 0; k < kMlpH; k++) {
    float z = theta[kMlpOffB1 + k];
#pragma unroll
    for (int c = 0; c < 4; c++) z = z + x[c] * w1q[c * kMlpH + k];
    h1[k] = fast_tanh(z);
  }
  __syncthreads();
  float y = theta[kMlpOffB3];
#pragma unroll
  for (int n = 0; n < kMlpH; n++) {
    float z = theta[kMlpOffB2 + n];
#pragma unroll
    for (int k = 0; k < kMlpH; k++) z = z + bf16_round(h1[k]) * w2s[k * kMlpH + n];
    h2[n] = fast_tanh(z);
    y = y + h2[n] * w3q[n];
  }
  const float r = valid ? y - t : 0.0f;
  const float d3 = 2.0f * w * r * inv_b;
  // phase 1: dW2 = h1q^T d2, db2 = sum d2
#pragma unroll
  for (int k = 0; k < kMlpH; k++) A[tid * 65 + k] = bf16_round(h1[k]);
  float d2[kMlpH];
#pragma unroll
  for (int n = 0; n < kMlpH; n++) {
    d2[n] = d3 * w3q[n] * (1.0f - h2[n] * h2[n]);
    Bm[tid * 65 + n] = d2[n];
  }
#pragma unroll
  for (int c = 0; c < 4; c++) xs[tid * 4 + c] = x[c];
  d3s[tid] = d3;
  __syncthreads();
  float *gp = gpart + (size_t)blockIdx.x * kMlpNParam;
  for (int e = tid; e < kMlpH * kMlpH; e += kGradBlock) {
    const int k = e >> 6, n = e & 63;
    float acc = 0.0f;
    for (int q = 0; q < kGradBlock; q++) acc = acc + A[q * 65 + k] * Bm[q * 65 + n];
    gp[kMlpOffW2 + e] = acc;
  }
  if (tid < kMlpH) {
    float acc = 0.0f;
    for (int q = 0; q < kGradBlock; q++) acc = acc + Bm[q * 65 + tid];
    gp[kMlpOffB2 + tid] = acc;
  }
  // dh1 = w2q d2, d1 = dh1 (1 - h1^2)
  float d1[kMlpH];
#pragma unroll
  for (int k = 0; k < kMlpH; k++) {
    float acc = 0.0f;
#pragma unroll
    for (int n = 0; n < kMlpH; n++) acc = acc + w2s[k * kMlpH + n] * d2[n];
    d1[k] = acc * (1.0f - h1[k] * h1[k]);
  }
  __syncthreads();
  // phase 2: dW3 = h2^T d3, db1 = sum d1, dW1 = x^T d1, db3 = sum d3
#pragma unroll
  for (int n = 0; n < kMlpH; n++) A[tid * 65 + n] = h2[n];
#pragma unroll
  for (int k = 0; k < kMlpH; k++) Bm[tid * 65 + k] = d1[k];
  __syncthreads();
  if (tid < kMlpH) {
    float acc = 0.0f;
    for (int q = 0; q < kGradBlock; q++) acc = acc + A[q * 65 + tid] * d3s[q];
    gp[kMlpOffW3 + tid] = acc;
  } else {
    const int k = tid - kMlpH;
    float acc = 0.0f;
    for (int q = 0; q < kGradBlock; q++) acc = acc + Bm[q * 65 + k];
    gp[kMlpOffB1 + k] = acc;
  }
  for (int e = tid; e < 4 * kMlpH; e += kGradBlock) {
    const int c = e >> 6, k = e & 63;
    float acc = 0.0f;
    for (int q = 0; q < kGradBlock; q++) acc = acc + xs[q * 4 + c] * Bm[q * 65 + k];
    gp[kMlpOffW1 + e] = acc;
  }
  // loss and db3 through the now-free d3s / A scratch
  __syncthreads();
  A[tid] = w * r * r;
  __syncthreads();
  if (tid == 0) {
    float acc = 0.0f, l = 0.0f;
    for (int q = 0; q < kGradBlock; q++) {
      acc = acc + d3s[q];
      l = l + A[q];
    }
    gp[kMlpOffB3] = acc;
    lpart[blockIdx.x] = l;
  }
}

// wa: the inference's MFMA A operands, w2a (4096 BF16) then w1a (1024 BF16,
// the lane group 3 slots left zero), layouts as documented at MlpDev; w3q is
// followed by the inference's scaled biases (MlpDev::bs).
__device__ __forceinline__ void store_quantized(int p, float w, float *w1q, float *w3q,
                                                unsigned short *wa) {
  if (p >= kMlpOffW1 && p < kMlpOffB1) {
    w1q[p - kMlpOffW1] = bf16_round(w);
    const int e = (p - kMlpOffW1) >> 6, n = (p - kMlpOffW1) & 63;
    unsigned short *w1a = wa + kMlpW2aShorts;
#pragma unroll
    for (int q = 0; q < 3; q++) w1a[((n >> 4) * 64 + 16 * q + (n & 15)) * 4 + e] = bf16_bits(w);
  } else if (p >= kMlpOffB1 && p < kMlpOffW2) {
    w3q[kMlpH + (p - kMlpOffB1)] = w * kMlpExp2Scale;
  } else if (p >= kMlpOffB2 && p < kMlpOffW3) {
    w3q[2 * kMlpH + (p - kMlpOffB2)] = w * kMlpExp2Scale;
  } else if (p >= kMlpOffW3 && p < kMlpOffB3) {
    w3q[p - kMlpOffW3] = bf16_round(w);
  } else if (p >= kMlpOffW2 && p < kMlpOffB2) {
    // w2[k][n]: k-slot e of lane group q in block kb as MlpDev documents
    const int k = (p - kMlpOffW2) >> 6, n = (p - kMlpOffW2) & 63;
    const int kb = k >> 5, kk = k & 31;
    const int e = 4 * (kk >> 4) + (kk & 3), lane = 16 * ((kk & 15) >> 2) + (n & 15);
    wa[(((n >> 4) * 2 + kb) * 64 + lane) * 8 + e] = bf16_bits(w);
  }
}

// Adam over the summed partials (bias-corrected, step t >= 1); also the
// plain "sum the partials" for gradient checks (lr = 0 and out != null).
__global__ void __launch_bounds__(kAdamBlock)
k_mlp_adam(float *__restrict__ theta, float *__restrict__ mo, float *__restrict__ ve,
           const float *__restrict__ gpart, int nblk, int t, float lr, float *__restrict__ w1q,
           float *__restrict__ w3q, unsigned short *__restrict__ wa, float *__restrict__ gout) {
  const int p = blockIdx.x * kAdamBlock + threadIdx.x;
  if (p >= kMlpNParam) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; b++) g = g + gpart[(size_t)b * kMlpNParam + p];
  if (gout) {
    gout[p] = g;
    return;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * mo[p] + (1.0f - b1) * g;
  const float v = b2 * ve[p] + (1.0f - b2) * g * g;
  mo[p] = m;
  ve[p] = v;
  const float mh = m / (1.0f - powf(b1, (float)t));
  const float vh = v / (1.0f - powf(b2, (float)t));
  const float w = theta[p] - lr * mh / (sqrtf(vh) + eps);
  theta[p] = w;
  store_quantized(p, w, w1q, w3q, wa);
}

__global__ void __launch_bounds__(kAdamBlock)
k_mlp_quantize(const float *__restrict__ theta, float *__restrict__ w1q, float *__restrict__ w3q,
               unsigned short *__restrict__ wa) {
  const int p = blockIdx.x * kAdamBlock + threadIdx.x;
  if (p >= kMlpNParam) return;
  store_quantized(p, theta[p], w1q, w3q, wa);
}

// Inference on raw inputs (log10 rho, log10 e, abar, zbar as floats):
// normalised output and the in-box flag.  Whole blocks run the wave-wide
// forward; lanes past n use zero inputs and store nothing.
__global__ void __launch_bounds__(kFwdBlock)
k_mlp_forward(MlpDev m, long long n, const float *__restrict__ raw, float *__restrict__ y,
              int *__restrict__ inbox) {
  __shared__ unsigned int lds[(kFwdBlock / 64) * kMlpLdsWaveDw];
  const long long i = (long long)blockIdx.x * kFwdBlock + threadIdx.x;
  const bool valid = i < n;
  float x[4];
  bool ok = valid;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    x[c] = valid ? (raw[4 * i + c] - m.in_off[c]) * m.in_scale[c] : 0.0f;
    ok = ok && fabsf(x[c]) <= m.box;
  }
  const float out = mlp_forward_wave(m, x, lds + (threadIdx.x >> 6) * kMlpLdsWaveDw);
  if (valid) {
    y[i] = out;
    if (inbox) inbox[i] = ok ? 1 : 0;
  }
}

// Batched T(rho, e) with a guess mode (GUESS_WARM / GUESS_MLP /
// GUESS_BRACKET); fallback flags per point.
__global__ void __launch_bounds__(kFwdBlock)
k_helm_rhoe_guess(HelmDev h, MlpDev m, int mode, long long n, const double *__restrict__ rho,
                  const double *__restrict__ e, const double *__restrict__ abar,
                  const double *__restrict__ zbar, double *__restrict__ T,
                  double *__restrict__ P, double *__restrict__ cs, int ret,
                  int *__restrict__ iters, int *__restrict__ fb) {
  __shared__ unsigned int lds[(kFwdBlock / 64) * kMlpLdsWaveDw];
  const long long i = (long long)blockIdx.x * kFwdBlock + threadIdx.x;
  const bool valid = i < n;
  double tg = -1.0;
  if (mode == GUESS_MLP) {
    float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool ok = false;
    if (valid) mlp_inputs(m, rho[i], e[i], abar[i], zbar[i], x, ok);
    const float y = mlp_forward_wave(m, x, lds + (threadIdx.x >> 6) * kMlpLdsWaveDw);
    if (ok) tg = exp10((double)(y * m.out_scale + m.out_off));
  }
  if (!valid) return;
  double t = T[i], p, s;
  int f;
  const int it = helm_rhoe_guess(h, mode, tg, m.resid_tol, rho[i], e[i], abar[i], zbar[i], t, p,
                                 s, ret != 0, f);
  helm_pin(p);
  helm_pin(s);
  T[i] = t;
  if (P) P[i] = p;
  if (cs) cs[i] = s;
  if (iters) iters[i] = it;
  if (fb) fb[i] = f;
}

inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_mlp_data(const HelmDev &h, long long n, unsigned long long seed, const double box[8],
                     float *raw, float *label, float *wt, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_data, dim3(nblocks(n, 256)), dim3(256), 0, st, h, n, seed, box[0],
                     box[1], box[2], box[3], box[4], box[5], box[6], box[7], raw, label, wt);
}
void launch_mlp_normalize(const MlpDev &m, long long n, const float *raw, const float *label,
                          float *xn, float *tn, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_normalize, dim3(nblocks(n, 256)), dim3(256), 0, st, m, n, raw, label,
                     xn, tn);
}
int launch_mlp_grad(const float *theta, const float *w1q, const float *w3q, const float *xn,
                    const float *tn, const float *wt, long long n0, long long nb, float *gpart,
                    float *lpart, hipStream_t st) {
  const int nblk = (int)nblocks(nb, kGradBlock);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(kGradBlock), 0, st, theta, w1q, w3q, xn, tn, wt,
                     n0, nb, 1.0f / (float)nb, gpart, lpart);
  return nblk;
}
void launch_mlp_adam(float *theta, float *m, float *v, const float *gpart, int nblk, int t,
                     float lr, float *w1q, float *w3q, unsigned short *wa, float *gout,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblocks(kMlpNParam, kAdamBlock)), dim3(kAdamBlock), 0, st,
                     theta, m, v, gpart, nblk, t, lr, w1q, w3q, wa, gout);
}
void launch_mlp_quantize(const float *theta, float *w1q, float *w3q, unsigned short *wa,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_quantize, dim3(nblocks(kMlpNParam, kAdamBlock)), dim3(kAdamBlock), 0,
                     st, theta, w1q, w3q, wa);
}
void launch_mlp_forward(const MlpDev &m, long long n, const float *raw, float *y, int *inbox,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_forward, dim3(nblocks(n, kFwdBlock)), dim3(kFwdBlock), 0, st, m, n, raw,
                     y, inbox);
}
void launch_helm_rhoe_guess(const HelmDev &h, const MlpDev &m, int mode, long long n,
                            const double *rho, const double *e, const double *abar,
                            const double *zbar, double *T, double *P, double *cs, int ret,
                            int *iters, int *fb, hipStream_t st) {
  hipLaunchKernelGGL(k_helm_rhoe_guess, dim3(nblocks(n, kFwdBlock)), dim3(kFwdBlock), 0, st, h, m,
                     mode, n, rho, e, abar, zbar, T, P, cs, ret, iters, fb);
}

}  // namespace phlx
