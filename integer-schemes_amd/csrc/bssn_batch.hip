// Batched block form: auxiliary kernels (queue reset, per-block periodic refill, halo copies,
// device time) and the dispatch of the batched stage kernels (bssn_batch.h).
#include <hip/hip_runtime.h>

#include "bssn_entries.h"
#include "bssn_hip.h"
#include "bssn_kernels.h"
#include "bssn_params.h"

namespace bssn {

hipError_t launch_fused_batch(const StageArgs& a, const BatchArgs& b, int grid, int ord, bool ko,
                              int mode, hipStream_t st) {
  return ord == 4   ? launch_batch_o4(a, b, grid, ko, mode, st)
         : ord == 6 ? launch_batch_o6(a, b, grid, ko, mode, st)
                    : launch_batch_o8(a, b, grid, ko, mode, st);
}

// Queue counters are cleared by a kernel (not a memset node) so that a captured step replays
// correctly.
__global__ void k_batch_reset(int* ctr) {
  if (threadIdx.x < BATCH_CHUNKS) ctr[threadIdx.x] = 0;
}

hipError_t launch_batch_reset(int* ctr, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, ctr);
  return hipGetLastError();
}

// Periodic ghost refill of every block (each block periodic on its own interior), one 64-lane
// workgroup per brick of the ghost list (the bricks of the padded block that hold a ghost
// point); a ghost point takes the value at its coordinates wrapped into the interior, which is
// what the slab copies x, then y, then z of 3D/src/bssn3d/evolve.py:65-73 leave there.
__global__ __launch_bounds__(64) void k_batch_periodic(double* __restrict__ pool,
                                                       const BlockDesc* __restrict__ blks,
                                                       BrickList ghost) {
  const int w = (int)blockIdx.x;
  const int ch = w & (BATCH_CHUNKS - 1);
  const int i = ghost.cb[ch] + (w >> 3);
  if (i >= ghost.cb[ch + 1]) return;
  const BrickItem item = ghost.items[i];
  const BlockDesc blk = blks[item.blk];
  const int lane = (int)threadIdx.x;
  const int c[3] = {(item.org & 1023) + brick_lx(lane), ((item.org >> 10) & 1023) + brick_ly(lane),
                    ((item.org >> 20) & 1023) + brick_lz(lane)};
  bool ghost_pt = false;
  int s[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (c[a] >= blk.S[a]) return;
    const int N = blk.S[a] - 2 * blk.ng;
    int v = (c[a] - blk.ng) % N;
    if (c[a] < blk.ng || c[a] >= blk.S[a] - blk.ng) ghost_pt = true;
    v = v < 0 ? v + N : v;
    s[a] = blk.ng + v;
  }
  if (!ghost_pt) return;
  const long P = (long)blk.S[0] * blk.S[1] * blk.S[2];
  const long dst = ((long)c[0] * blk.S[1] + c[1]) * blk.S[2] + c[2];
  const long src = ((long)s[0] * blk.S[1] + s[1]) * blk.S[2] + s[2];
  double* u = pool + blk.off;
  // all 24 loads before the stores (a ghost destination is never an interior source), so they
  // are in flight together instead of each waiting behind the previous store
  double v[BSSN_NF];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) v[f] = u[(long)f * P + src];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) u[(long)f * P + dst] = v[f];
}

hipError_t launch_batch_periodic(double* pool, const BlockDesc* blk, const BrickList& ghost,
                                 int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_periodic, dim3(grid), dim3(64), 0, st, pool, blk, ghost);
  return hipGetLastError();
}

// Ghosts held for a step: every ghost point of every block is copied from U into the two stage
// buffers (all 24 fields), so stages whose ghosts are not refilled by a halo map read the
// ghosts of the step's base state (Dirichlet data from the last sync).
__global__ __launch_bounds__(64) void k_batch_hold(const double* __restrict__ U,
                                                   double* __restrict__ Ya,
                                                   double* __restrict__ Yb,
                                                   const BlockDesc* __restrict__ blks,
                                                   BrickList ghost) {
  const int w = (int)blockIdx.x;
  const int ch = w & (BATCH_CHUNKS - 1);
  const int i = ghost.cb[ch] + (w >> 3);
  if (i >= ghost.cb[ch + 1]) return;
  const BrickItem item = ghost.items[i];
  const BlockDesc blk = blks[item.blk];
  const int lane = (int)threadIdx.x;
  const int c[3] = {(item.org & 1023) + brick_lx(lane), ((item.org >> 10) & 1023) + brick_ly(lane),
                    ((item.org >> 20) & 1023) + brick_lz(lane)};
  bool ghost_pt = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (c[a] >= blk.S[a]) return;
    if (c[a] < blk.ng || c[a] >= blk.S[a] - blk.ng) ghost_pt = true;
  }
  if (!ghost_pt) return;
  const long P = (long)blk.S[0] * blk.S[1] * blk.S[2];
  const long p = blk.off + ((long)c[0] * blk.S[1] + c[1]) * blk.S[2] + c[2];
#pragma unroll 4
  for (int f = 0; f < BSSN_NF; ++f) {
    const double v = U[p + (long)f * P];
    Ya[p + (long)f * P] = v;
    Yb[p + (long)f * P] = v;
  }
}

hipError_t launch_batch_hold(const double* U, double* Ya, double* Yb, const BlockDesc* blk,
                             const BrickList& ghost, int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_hold, dim3(grid), dim3(64), 0, st, U, Ya, Yb, blk, ghost);
  return hipGetLastError();
}

// Ghost exchange between blocks: entry e copies all 24 fields of one point,
// pool[dst + f * dst_stride] = pool[src + f * src_stride]. Sources are interior points, which
// no entry writes, so the order of the copies does not matter.
// An entry that would reach outside the pool (pool_n doubles) is skipped.
__global__ __launch_bounds__(256) void k_halo_copy(double* __restrict__ pool,
                                                   const bssn_halo_copy* __restrict__ map,
                                                   long n, long long pool_n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const bssn_halo_copy m = map[e];
  if (!halo_entry_ok(m, pool_n)) return;
  // all 24 loads before the stores (an entry's interior source is never a ghost destination),
  // so they are in flight together instead of each waiting behind the previous store
  double v[BSSN_NF];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) v[f] = pool[m.src + (long)f * m.src_stride];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) pool[m.dst + (long)f * m.dst_stride] = v[f];
}

hipError_t launch_halo_copy(double* pool, long long pool_n, const void* map, long n,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_halo_copy, dim3((unsigned)blocks), dim3(256), 0, st, pool,
                     (const bssn_halo_copy*)map, n, pool_n);
  return hipGetLastError();
}

// device-resident time of a captured step: t += dt after the step's last stage
__global__ void k_time_advance(double* t, double dt) {
  if (threadIdx.x == 0) *t = *t + dt;
}

hipError_t launch_time_advance(double* tdev, double dt, hipStream_t st) {
  hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(64), 0, st, tdev, dt);
  return hipGetLastError();
}

// t += dt inside a scalar pack held in device memory
__global__ void k_pack_time_advance(double* S) {
  if (threadIdx.x == 0) S[PK_T] = S[PK_T] + S[PK_DT];
}

hipError_t launch_pack_time_advance(double* S, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_time_advance, dim3(1), dim3(64), 0, st, S);
  return hipGetLastError();
}

}  // namespace bssn
