// Fused stage kernels of FD order 4 (one compile unit per order; kernels in bssn_fused.h).
#include "bssn_fused.h"

namespace bssn {

hipError_t launch_fused_o4(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st) {
  return launch_fused_ord<4>(a, d, w, sc, ko, mode, st);
}

hipError_t launch_fused(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                        int ord, bool ko, int mode, hipStream_t st) {
  if (a.sdev)
    return ord == 4   ? launch_fused_o4s(a, d, w, sc, ko, mode, st)
           : ord == 6 ? launch_fused_o6s(a, d, w, sc, ko, mode, st)
                      : launch_fused_o8s(a, d, w, sc, ko, mode, st);
  return ord == 4   ? launch_fused_o4(a, d, w, sc, ko, mode, st)
         : ord == 6 ? launch_fused_o6(a, d, w, sc, ko, mode, st)
                    : launch_fused_o8(a, d, w, sc, ko, mode, st);
}

}  // namespace bssn
