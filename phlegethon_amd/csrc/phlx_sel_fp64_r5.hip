// phlx_sel_fp64_r5.hip -- kernel instances for FP64, LIM5TH
// (one compilation unit per family so the library builds in parallel).
#include "phlx_kernels.hpp"

namespace phlx {

bool select_fp64_r5(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  return select_prec<false, true>(sd, nas, mhd, helm, fasteos, t);
}

}  // namespace phlx
