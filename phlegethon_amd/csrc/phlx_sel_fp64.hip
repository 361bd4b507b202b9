// phlx_sel_fp64.hip -- kernel instances for FP64, 2nd-order reconstructions
// (one compilation unit per family so the library builds in parallel).
#include "phlx_kernels.hpp"

namespace phlx {

bool select_fp64(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  return select_prec<false, false>(sd, nas, mhd, helm, fasteos, t);
}

}  // namespace phlx
