// Batched stage kernels of FD order 8 (one compile unit per order; kernels in bssn_batch.h).
#include "bssn_batch.h"

namespace bssn {

hipError_t launch_batch_o8(const StageArgs& a, const BatchArgs& b, int grid, bool ko, int mode,
                            hipStream_t st) {
  return launch_batch_ord<8>(a, b, grid, ko, mode, st);
}

}  // namespace bssn
