// Fused stage kernels of FD order 8 (one compile unit per order; kernels in bssn_fused.h).
#include "bssn_fused.h"

namespace bssn {

hipError_t launch_fused_o8(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st) {
  return launch_fused_ord<8>(a, d, w, sc, ko, mode, st);
}

}  // namespace bssn
