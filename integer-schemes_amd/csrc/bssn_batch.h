// Batched form of the fused RHS / RK stage kernels: every block of a block-structured
// hierarchy (blocks of different extents and spacings in one pool) in one launch.
//
// Replaces, on the reference side, the slot-batched step of the block-AMR solver
// (2D/src/mcs2d/amr/evolve.py:659-775): there every level is a fixed-shape slot array and the
// RHS is vmapped over all slots, inactive ones included (:710-717). Here the work is a list of
// bricks over the blocks that exist, of any extent, and each brick is evaluated by the
// same wave program as the single-grid kernel (bssn_fused.h), so a block's result is
// bit-identical to running the single-grid kernel on that block alone.
//
// Scheduling: the brick list is split into one chunk per XCD. Either one workgroup per brick
// (workgroup w -> chunk w mod 8, matching the round-robin placement of workgroups on XCDs), or
// a persistent grid whose waves pull bricks from their XCD's queue and then from the others'
// until every queue is drained; each wave takes one brick per atomic, so every wave exits once
// the counters pass the chunk ends.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

#include "bssn_algebra.h"
#include "bssn_brick.h"
#include "bssn_hip.h"
#include "bssn_kernels.h"
#include "bssn_params.h"
#include "bssn_stage.h"

namespace bssn {


__device__ __forceinline__ const double* shifted(const double* p, long off) {
  return p ? p + off : p;
}
__device__ __forceinline__ double* shifted(double* p, long off) { return p ? p + off : p; }

// next brick of this wave: -1 when every chunk it may serve is drained
__device__ __forceinline__ int next_item(const BatchArgs& b, int* chunk, int* round) {
  if (!b.persistent) {
    if (*round) return -1;
    *round = 1;
    const int w = (int)blockIdx.x;
    const int c = w & (BATCH_CHUNKS - 1);
    const int i = b.list.cb[c] + (w >> 3);
    return i < b.list.cb[c + 1] ? i : -1;
  }
  while (*round < BATCH_CHUNKS) {
    const int c = (*chunk + *round) & (BATCH_CHUNKS - 1);
    int k = 0;
    if (threadIdx.x == 0) k = atomicAdd(&b.ctr[c], 1);
    k = __builtin_amdgcn_readfirstlane(k);
    const int i = b.list.cb[c] + k;
    if (i < b.list.cb[c + 1]) return i;
    ++*round;
  }
  return -1;
}

// one brick of one block, by the single-grid wave program
template <int ORD, bool KO, int MODE, int KIND>
__device__ __forceinline__ void batch_brick(double* lds, const StageArgs& args,
                                            const BatchArgs& b, int it, int lane) {
  using G = Geo<ORD, KO>;
  const BrickItem item = b.list.items[it];
  const BlockDesc blk = b.blk[item.blk];
  Dims d;
  d.ng = blk.ng;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d.S[a] = blk.S[a];
    d.lo[a] = b.interior ? blk.ng : 0;
    d.n[a] = b.interior ? blk.S[a] - 2 * blk.ng : blk.S[a];
  }
  d.P = (long)blk.S[0] * blk.S[1] * blk.S[2];
  const int ox = item.org & 1023, oy = (item.org >> 10) & 1023, oz = (item.org >> 20) & 1023;
  Scales w = make_scales(blk.h, ORD, b.sigma);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    w.s1[a] = uniform(w.s1[a]);
    w.s2[a] = uniform(w.s2[a]);
    w.sm[a] = uniform(w.sm[a]);
    w.sko[a] = uniform(w.sko[a]);
  }
  Scal sc = with_ssl(b.base, args.tdev, args.toff);
  sc.cahd = uniform(cahd_factor(b.cahd_c, blk.h[0], b.dt));
  StageArgs a = args;
  a.Yin = shifted(args.Yin, blk.off);
  a.U = shifted(args.U, blk.off);
  a.Y = shifted(args.Y, blk.off);
  a.K1 = shifted(args.K1, blk.off);
  a.FP = shifted(args.FP, blk.off);
  a.Uout = shifted(args.Uout, blk.off);
  using Src = WaveSrc<ORD, KO, MODE, RhsOrder, fused_depth<KIND, ORD>()>;
  if constexpr (MODE == LINK) {
    Src src(lds, args.Yin, blk.off, b.link + 2 * blk.moff, b.pool_n, w, d, ox, oy, oz, lane);
    StageOut<KIND> out(a, d.P, src.gc, lane_valid<G>(d, ox, oy, oz, lane), true);
    bssn_rhs_point<KO>(src, sc, out);
  } else {
    Src src(lds, a.Yin, w, d, ox, oy, oz, lane);
    StageOut<KIND> out(a, d.P, src.gc, lane_valid<G>(d, ox, oy, oz, lane), true);
    bssn_rhs_point<KO>(src, sc, out);
  }
}

template <int ORD, bool KO, int MODE, bool PERSIST, int KIND>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(fused_waves<KIND>(), fused_waves<KIND>())))
void k_fused_batch(StageArgs args, BatchArgs b) {
  using G = Geo<ORD, KO>;
  __shared__ double lds[G::VOL];
  const int lane = (int)threadIdx.x;
  int chunk = (int)blockIdx.x & (BATCH_CHUNKS - 1), round = 0;
  if (!PERSIST) {
    const int it = next_item(b, &chunk, &round);
    if (it >= 0) batch_brick<ORD, KO, MODE, KIND>(lds, args, b, it, lane);
    return;
  }
#pragma unroll 1
  for (int it = next_item(b, &chunk, &round); it >= 0; it = next_item(b, &chunk, &round)) {
    batch_brick<ORD, KO, MODE, KIND>(lds, args, b, it, lane);
    // the next brick restages the box: keep its writes behind this brick's reads
    __builtin_amdgcn_wave_barrier();
  }
}

template <int ORD, bool KO, int MODE, int KIND>
static hipError_t launch_batch_t(const StageArgs& a, const BatchArgs& b, int grid,
                                 hipStream_t st) {
  using G = Geo<ORD, KO>;
  if (b.persistent)
    hipLaunchKernelGGL((k_fused_batch<ORD, KO, MODE, true, KIND>), dim3(grid), dim3(G::NT), 0,
                       st, a, b);
  else
    hipLaunchKernelGGL((k_fused_batch<ORD, KO, MODE, false, KIND>), dim3(grid), dim3(G::NT), 0,
                       st, a, b);
  return hipGetLastError();
}

// every (KO, MODE, kind) instance of one FD order (RHS and RK4 kinds)
template <int ORD>
static hipError_t launch_batch_ord(const StageArgs& a, const BatchArgs& b, int grid, bool ko,
                                   int mode, hipStream_t st) {
  return with_kind<false>(a.kind, [&](auto k) -> hipError_t {
    constexpr int K = decltype(k)::value;
    if (ko)
      return mode == WRAP   ? launch_batch_t<ORD, true, WRAP, K>(a, b, grid, st)
             : mode == LINK ? launch_batch_t<ORD, true, LINK, K>(a, b, grid, st)
                            : launch_batch_t<ORD, true, CLAMP, K>(a, b, grid, st);
    return mode == WRAP   ? launch_batch_t<ORD, false, WRAP, K>(a, b, grid, st)
           : mode == LINK ? launch_batch_t<ORD, false, LINK, K>(a, b, grid, st)
                          : launch_batch_t<ORD, false, CLAMP, K>(a, b, grid, st);
  });
}

}  // namespace bssn
