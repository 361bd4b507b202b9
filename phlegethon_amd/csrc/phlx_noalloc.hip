// phlx_noalloc.hip -- compute_hyperbolic_dt on caller-owned device buffers,
// allocating nothing: for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops that forbid hipMalloc).  The
// per-block partial maxima live in a file-scope device array; its address
// is taken once per device with hipGetSymbolAddress on the entry point's
// first call there.
#include <cstdio>
#include <mutex>

#include "phlx_kernels.hpp"
#include "../../include/phlx.h"

namespace phlx {

// partial maxima of the signal speed, one per block of the reduction grid
// (a grid-stride launch of at most this many blocks covers any grid)
constexpr int kNoAllocBlocks = 2048;
static __device__ __align__(256) double scratch[kNoAllocBlocks];

constexpr int kNoAllocMaxDevices = 64;
static double *g_scratch_dev[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (the array's host-side
// name is not a device address).
static int scratch_address(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return PHLX_E_HIP;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return PHLX_E_CONFIG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PHLX_E_HIP;
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return PHLX_OK;
}

// Signal speed sum_d (|v_d| + c_f) / dx_d of every cell (cell_speed, the
// arithmetic of the library's CFL pass), grid-stride; each block's maximum
// into part[blockIdx.x].
template <int SD, bool MHD>
__global__ void __launch_bounds__(256) k_cfl_partials(Grid g, double *part) {
  constexpr int NL = SD + 2;
  double s = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N; c += stride) {
    double p[NL], b[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < NL; v++) p[v] = g.P[v * g.N + c];
    if (MHD) {
#pragma unroll
      for (int v = 0; v < SD; v++) b[v] = g.bcc[v * g.N + c];
    }
    s = fmax(s, cell_speed<SD, MHD, false>(g, p, b, 0.0));
  }
  __shared__ double wmax[4];
  s = wave_max(s);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// dt = cfl / max, smax = max of the partials (the library's step-end rule).
__global__ void __launch_bounds__(1024) k_cfl_finish(const double *part, int n, double cfl,
                                                     double *out) {
  const double m = reduce_partials(part, n);
  if (threadIdx.x == 0) {
    out[0] = cfl / m;
    out[1] = m;
  }
}

}  // namespace phlx

using namespace phlx;

// largest grid accepted: 2^33 cells (a 64 GiB field, beyond one GPU's memory
// for the nv >= 4 fields the pass reads)
static constexpr long long kNoAllocMaxCells = 1LL << 33;

extern "C" int phlx_compute_dt_noalloc(const phlx_config *cfg, const double *prim,
                                       const double *bcc, double *dt_smax, void *stream) {
  if (!cfg || !prim || !dt_smax) return PHLX_E_ARG;
  if (cfg->sdims != 2 && cfg->sdims != 3) return PHLX_E_CONFIG;
  if (cfg->eos != PHLX_EOS_IDEAL) return PHLX_E_CONFIG;   // the tabulated EoS needs a table
  if (!(cfg->gamma > 1.0) || !(cfg->cfl > 0.0)) return PHLX_E_CONFIG;
  const long long n1 = cfg->nx1, n2 = cfg->nx2, n3 = cfg->sdims == 3 ? cfg->nx3 : 1;
  if (n1 < 1 || n2 < 1 || n3 < 1) return PHLX_E_ARG;
  const long long N = n1 * n2 * n3;
  if (N > kNoAllocMaxCells) {
    std::fprintf(stderr, "phlx_compute_dt_noalloc: %lld cells exceed the maximum of %lld\n",
                 N, kNoAllocMaxCells);
    return PHLX_E_ARG;
  }
  const bool mhd = cfg->flux == PHLX_HLLD_FLUX || cfg->flux == PHLX_LHLLD_FLUX;
  if (mhd && !bcc) return PHLX_E_ARG;
  double *part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  Grid g{};
  g.n1 = (int)n1;
  g.n2 = (int)n2;
  g.n3 = (int)n3;
  g.sd = cfg->sdims;
  g.N = N;
  g.gm = cfg->gamma;
  g.inv_dx[0] = 1.0 / ((cfg->x1u - cfg->x1l) / (double)cfg->nx1);
  g.inv_dx[1] = 1.0 / ((cfg->x2u - cfg->x2l) / (double)cfg->nx2);
  g.inv_dx[2] = cfg->sdims == 3 ? 1.0 / ((cfg->x3u - cfg->x3l) / (double)cfg->nx3) : 1.0;
  g.P = const_cast<double *>(prim);
  g.bcc = const_cast<double *>(bcc);
  const long long want = (N + 255) / 256;
  const int blocks = (int)(want < kNoAllocBlocks ? want : kNoAllocBlocks);
  hipStream_t st = (hipStream_t)stream;
  if (cfg->sdims == 3) {
    if (mhd) k_cfl_partials<3, true><<<blocks, 256, 0, st>>>(g, part);
    else k_cfl_partials<3, false><<<blocks, 256, 0, st>>>(g, part);
  } else {
    if (mhd) k_cfl_partials<2, true><<<blocks, 256, 0, st>>>(g, part);
    else k_cfl_partials<2, false><<<blocks, 256, 0, st>>>(g, part);
  }
  k_cfl_finish<<<1, 1024, 0, st>>>(part, blocks, cfg->cfl, dt_smax);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::fprintf(stderr, "phlx_compute_dt_noalloc: launch failed: %s\n", hipGetErrorString(e));
    return PHLX_E_HIP;
  }
  return PHLX_OK;
}
