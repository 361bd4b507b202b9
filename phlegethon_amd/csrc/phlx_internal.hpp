// phlx_internal.hpp -- shared declarations between the kernels and the C-ABI.
#pragma once

#include "phlx_device.hpp"

namespace phlx {

// Launchers of one template instance (space dims, advected scalars, MHD).
struct KernelTable {
  void (*flux)(const Grid &, int, hipStream_t);
  void (*corner)(const Grid &, hipStream_t);
  void (*faceb)(const Grid &, int, double, double, double, hipStream_t);
  void (*update)(const Grid &, int, int, double, double, double, hipStream_t);
  void (*cfl)(const Grid &, hipStream_t);
  void (*init)(const Grid &, hipStream_t);
  void (*divb)(const Grid &, hipStream_t);
};

bool select_kernels(int sd, int nas, bool mhd, KernelTable &t);
void launch_step_begin(Scalars *sc, hipStream_t st);
void launch_step_end(const Grid &g, int advance, hipStream_t st);
void launch_reduce_to(const Grid &g, double *out, hipStream_t st);
int cell_grid_blocks(const Grid &g);
void launch_stream_copy(const void *src, void *dst, long long n16, hipStream_t st);

enum { KID_FLUX1 = 0, KID_FLUX2, KID_FLUX3, KID_CORNER, KID_FACEB, KID_UPDATE, KID_MISC, KID_COUNT };

}  // namespace phlx
