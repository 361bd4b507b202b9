// phlx_kernels.hip -- step-bracket, state-flag, bandwidth-probe and shock-flag
// launchers, and the kernel-table dispatch over the per-precision units.
#include "phlx_kernels.hpp"

namespace phlx {

// Device state flags (one thread): flags = (flags & keep) | set.
__global__ void k_flags(Scalars *sc, int keep, int set) { sc->flags = (sc->flags & keep) | set; }

// Step bracket kernels (one thread): time_loop's dt clip before the step
// and time/dt/step update after it (source/source.F90:6662-6720).
__global__ void k_step_begin(Scalars *sc) {
  // the loop runs while time < tmax (source/source.F90:6525)
  sc->active = (sc->time < sc->tmax) ? 1 : 0;
  if (!sc->active) return;
  if (sc->time + sc->dt > sc->tmax) sc->dt = sc->tmax - sc->time;
}
// dt = cfl / max of the partials; with advance: time += dt first, step++.
__global__ void __launch_bounds__(1024) k_step_end(Scalars *sc, const double *red, int n,
                                                   int advance) {
  if (advance && !sc->active) return;
  const double m = reduce_partials(red, n);
  if (threadIdx.x == 0) {
    if (advance) sc->time = sc->time + sc->dt;
    sc->smax = m;
    sc->dt = sc->cfl / m;
    if (advance) sc->step = sc->step + 1;
    sc->active = 1;
  }
}
__global__ void __launch_bounds__(1024) k_reduce_to(const double *red, int n, double *out) {
  const double m = reduce_partials(red, n);
  if (threadIdx.x == 0) *out = m;
}

// Plain vector copy used as the bandwidth reference (16 B per lane access).
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ src,
                                                     double2 *__restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

void launch_stream_copy(const void *src, void *dst, long long n16, hipStream_t st) {
  const int blocks = 256 * 8 * 4;
  k_stream_copy<<<blocks, 256, 0, st>>>((const double2 *)src, (double2 *)dst, n16);
}

void launch_shock_flags(const Grid &g, hipStream_t st) {
  const dim3 blk(64, 4, 1);
  if (g.sd == 3) {
    k_sf_detect<3><<<grid_plane(g.n1 + 4, g.n2 + 4, g.n3 + 4), blk, 0, st>>>(g);
    k_sf_dilate<3><<<grid_plane(g.n1 + 2, g.n2 + 2, g.n3 + 2), blk, 0, st>>>(g);
  } else {
    k_sf_detect<2><<<grid_plane(g.n1 + 4, g.n2 + 4, 1), blk, 0, st>>>(g);
    k_sf_dilate<2><<<grid_plane(g.n1 + 2, g.n2 + 2, 1), blk, 0, st>>>(g);
  }
}

bool select_kernels(int sd, int nas, bool mhd, bool helm, bool fasteos, bool mixed, bool r5,
                    KernelTable &t) {
  if (mixed)
    return r5 ? select_mixed_r5(sd, nas, mhd, helm, fasteos, t)
              : select_mixed(sd, nas, mhd, helm, fasteos, t);
  return r5 ? select_fp64_r5(sd, nas, mhd, helm, fasteos, t)
            : select_fp64(sd, nas, mhd, helm, fasteos, t);
}

void launch_step_begin(Scalars *sc, hipStream_t st) { k_step_begin<<<1, 1, 0, st>>>(sc); }
void launch_flags(Scalars *sc, int keep, int set, hipStream_t st) {
  k_flags<<<1, 1, 0, st>>>(sc, keep, set);
}
void launch_step_end(const Grid &g, int advance, hipStream_t st) {
  k_step_end<<<1, 1024, 0, st>>>(g.sc, g.red, g.nred, advance);
}
void launch_reduce_to(const Grid &g, double *out, hipStream_t st) {
  k_reduce_to<<<1, 1024, 0, st>>>(g.red, g.nred, out);
}
int cell_grid_blocks(const Grid &g) {
  const dim3 blk(64, 4, 1);
  const dim3 gr = grid_for(g.n1, g.n2, g.n3, blk);
  return (int)(gr.x * gr.y * gr.z);
}

}  // namespace phlx
