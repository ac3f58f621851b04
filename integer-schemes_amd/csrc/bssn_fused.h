// Fused BSSN right-hand side and fused RK4 / MSRK stage kernels for gfx950.
//
// One kernel evaluates, per output point, the 138 finite-difference derivatives (from
// wave-private LDS bricks, bssn_brick.h), the Kreiss-Oliger terms, the pointwise BSSN algebra
// (bssn_algebra.h) and the write-back: either the bare RHS, or an RK stage update
// (next stage input + running combination), with the algebraic enforcement of the last stage
// folded into the final store.
//
// Replaces, on the reference side:
//   bare RHS    : the M4 kernel bssn_rhs_fused behind the FFI handler Fused
//                 (3D/src/bssn3d/cuda/rhs_fused.cu:22-61, gen_fused.py:182-262)
//   RK4 stages  : BSSNEvolution.step (3D/src/bssn3d/evolve.py:125-132) with rhs() =
//                 periodic sync + RHS + KO (evolve.py:86-90) and enforce (evolve.py:93-119)
//   MSRK stages : BSSNEvolution._msrk2_step (evolve.py:139-153)
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

#include "bssn_algebra.h"
#include "bssn_brick.h"
#include "bssn_kernels.h"
#include "bssn_params.h"
#include "bssn_stage.h"

namespace bssn {

// SDEV: the scalars come from a pack in device memory (StageArgs::sdev). A separate
// instantiation, so that the host-pack kernels keep their scalars in the kernel arguments
// (a mutable copy of the arguments stays live in registers through the whole kernel).
template <int ORD, bool KO, int MODE, int KIND, bool SDEV>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(fused_waves<KIND>(), fused_waves<KIND>())))
void k_fused(StageArgs args, Dims d, Scales w, Scal sc) {
  using G = Geo<ORD, KO>;
  __shared__ double lds[G::VOL];
  int ox, oy, oz;
  if (!brick_origin<G>(d, &ox, &oy, &oz)) return;
  const int lane = (int)threadIdx.x;
  if constexpr (SDEV) args_from_pack<ORD>(args, w, sc);
  WaveSrc<ORD, KO, MODE, RhsOrder, fused_depth<KIND, ORD>()> src(lds, args.Yin, w, d, ox, oy, oz, lane);
  StageOut<KIND> out(args, d.P, src.gc, lane_valid<G>(d, ox, oy, oz, lane),
                     true);
  bssn_rhs_point<KO>(src, with_ssl(sc, args.tdev, args.toff), out);
}

template <int ORD, bool KO, int MODE, int KIND, bool SDEV>
static hipError_t launch_fused_t(const StageArgs& a, const Dims& d, const Scales& w,
                                 const Scal& sc, hipStream_t st) {
  using G = Geo<ORD, KO>;
  hipLaunchKernelGGL((k_fused<ORD, KO, MODE, KIND, SDEV>), dim3(brick_blocks<G>(d)),
                     dim3(G::NT), 0, st, a, d, w, sc);
  return hipGetLastError();
}

// every (KO, MODE, kind) instance of one FD order with the host pack
template <int ORD>
static hipError_t launch_fused_ord(const StageArgs& a, const Dims& d, const Scales& w,
                                   const Scal& sc, bool ko, int mode, hipStream_t st) {
  return with_kind<true>(a.kind, [&](auto k) -> hipError_t {
    constexpr int K = decltype(k)::value;
    if (ko) return mode == WRAP ? launch_fused_t<ORD, true, WRAP, K, false>(a, d, w, sc, st)
                                : launch_fused_t<ORD, true, CLAMP, K, false>(a, d, w, sc, st);
    return mode == WRAP ? launch_fused_t<ORD, false, WRAP, K, false>(a, d, w, sc, st)
                        : launch_fused_t<ORD, false, CLAMP, K, false>(a, d, w, sc, st);
  });
}

// the device-pack instances (bare RHS and RK4 kinds) of one FD order
template <int ORD>
static hipError_t launch_fused_ord_sdev(const StageArgs& a, const Dims& d, const Scales& w,
                                        const Scal& sc, bool ko, int mode, hipStream_t st) {
  return with_kind<false>(a.kind, [&](auto k) -> hipError_t {
    constexpr int K = decltype(k)::value;
    if (ko) return mode == WRAP ? launch_fused_t<ORD, true, WRAP, K, true>(a, d, w, sc, st)
                                : launch_fused_t<ORD, true, CLAMP, K, true>(a, d, w, sc, st);
    return mode == WRAP ? launch_fused_t<ORD, false, WRAP, K, true>(a, d, w, sc, st)
                        : launch_fused_t<ORD, false, CLAMP, K, true>(a, d, w, sc, st);
  });
}

}  // namespace bssn
