// phlx_sel_mixed_r5.hip -- kernel instances for mixed precision, LIM5TH
// (one compilation unit per family so the library builds in parallel).
#include "phlx_kernels.hpp"

namespace phlx {

bool select_mixed_r5(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  KernelTable base{}, r5{};
  if (!select_mixed(sd, nas, mhd, helm, fasteos, base)) return false;
  if (!select_prec<true, true>(sd, nas, mhd, helm, fasteos, r5)) return false;
  base.flux = r5.flux;   // LIM5TH sweeps; update, CT and EoS kernels shared
  base.sweep = nullptr;  // the fused sweep stages the 2nd-order stencil only
  base.update_swept = nullptr;
  base.update_swept_guess = nullptr;
  t = base;
  return true;
}

}  // namespace phlx
