// Fused RHS and RK4 stage kernels of FD order 8 reading their scalars from a pack in device
// memory (kernels in bssn_fused.h; a compile unit of their own, built in parallel).
#include "bssn_fused.h"

namespace bssn {

hipError_t launch_fused_o8s(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st) {
  return launch_fused_ord_sdev<8>(a, d, w, sc, ko, mode, st);
}

}  // namespace bssn
