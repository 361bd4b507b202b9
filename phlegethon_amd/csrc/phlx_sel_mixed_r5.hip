// phlx_sel_mixed_r5.hip -- kernel instances for mixed precision, LIM5TH
// (one compilation unit per family so the library builds in parallel).
#include "phlx_kernels.hpp"

namespace phlx {

bool select_mixed_r5(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  return select_prec<true, true>(sd, nas, mhd, helm, fasteos, t);
}

}  // namespace phlx
