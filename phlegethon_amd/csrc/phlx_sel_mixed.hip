// phlx_sel_mixed.hip -- kernel instances for mixed precision, 2nd-order reconstructions
// (one compilation unit per family so the library builds in parallel).
#include "phlx_kernels.hpp"

namespace phlx {

bool select_mixed(int sd, int nas, bool mhd, bool helm, bool fasteos, KernelTable &t) {
  return select_prec<true, false>(sd, nas, mhd, helm, fasteos, t);
}

}  // namespace phlx
