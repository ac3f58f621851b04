// LDS brick staging for the fused kernels.
//
// A workgroup of 256 threads owns a brick of BX x BY x BZ = 4 x 8 x 8 output points (one point
// per thread, z fastest so a wave covers one 8 x 8 (y, z) plane). Fields are staged one at a
// time into a (BX+2H)(BY+2H)(BZ+2H) LDS box, double-buffered so a single barrier per field
// separates the writes of field f+1 from the reads of field f-1. Every derivative of the
// staged field is then a plain stencil over LDS.
//
// Halo positions are mapped to the grid once per thread (not per field):
//   CLAMP  x -> min(max(x, 0), S-1)  per axis. The reference edge-pads each one-axis operator by
//          its own reach and forms mixed derivatives as d1_j(d1_i f)
//          (common/src/mcs_common/derivatives.py:78-88, derivative_bundle.py:41-46); clamping is
//          separable, so stencils over the clamped box equal those exactly, ghosts included.
//   WRAP   x -> ng + ((x - ng) mod N) per axis (N = S - 2 ng): the values a periodic ghost sync
//          (3D/src/bssn3d/evolve.py:65-73) would have put there, so no sync pass is needed
//          between RK stages.
#pragma once

#include <hip/hip_runtime.h>

#include "bssn_params.h"

namespace bssn {


template <int ORD, bool KO>
struct Geo {
  static constexpr int R = ORD / 2;                       // FD reach
  static constexpr int RK = KO ? (ORD + 2) / 2 : 0;       // KO reach (order ORD+2)
  static constexpr int H = R > RK ? R : RK;               // brick halo
  static constexpr int BX = 4, BY = 8, BZ = 8;
  static constexpr int NT = BX * BY * BZ;                 // threads per workgroup
  static constexpr int NX = BX + 2 * H, NY = BY + 2 * H, NZ = BZ + 2 * H;
  static constexpr int SXL = NY * NZ, SYL = NZ;           // LDS strides
  static constexpr int VOL = NX * NY * NZ;
  static constexpr int PER = (VOL + NT - 1) / NT;         // staged values per thread
  static constexpr int LVOL = PER * NT;                   // LDS slots per buffer (padded)
  static constexpr int NTAP = ORD + 1;
  static constexpr int NKO = ORD + 3;
};

// Map brick-local positions to flat grid offsets for this thread's staging slots.
template <class G, int MODE>
__device__ __forceinline__ void stage_offsets(const Dims& d, int ox, int oy, int oz, int* off) {
#pragma unroll
  for (int k = 0; k < G::PER; ++k) {
    const int e = (int)threadIdx.x + k * G::NT;
    const int ee = e < G::VOL ? e : G::VOL - 1;
    const int lz = ee % G::NZ;
    const int t = ee / G::NZ;
    const int ly = t % G::NY;
    const int lx = t / G::NY;
    int c[3] = {ox + lx - G::H, oy + ly - G::H, oz + lz - G::H};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (MODE == CLAMP) {
        c[a] = c[a] < 0 ? 0 : (c[a] > d.S[a] - 1 ? d.S[a] - 1 : c[a]);
      } else {
        const int N = d.S[a] - 2 * d.ng;
        int v = (c[a] - d.ng) % N;   // once per thread and slot, not per field
        v = v < 0 ? v + N : v;
        c[a] = d.ng + v;
      }
    }
    off[k] = (c[0] * d.S[1] + c[1]) * d.S[2] + c[2];
  }
}

// Source of fields and derivatives for bssn_rhs_point: one LDS brick per staged field.
template <int ORD, bool KO, int MODE>
struct BrickSrc {
  using G = Geo<ORD, KO>;
  double* lds;          // 2 * LVOL doubles
  const double* F;      // stencil source state
  const FDW& w;
  long P;
  int off[G::PER];      // staging offsets
  long gc;              // flat index of the thread's point (clamped into the grid)
  int c;                // LDS index of the thread's point
  int buf;

  __device__ __forceinline__ BrickSrc(double* lds_, const double* F_, const FDW& w_, const Dims& d,
                                      int ox, int oy, int oz, int tx, int ty, int tz)
      : lds(lds_), F(F_), w(w_), P(d.P), buf(1) {
    stage_offsets<G, MODE>(d, ox, oy, oz, off);
    c = ((tx + G::H) * G::NY + (ty + G::H)) * G::NZ + (tz + G::H);
    int p[3] = {ox + tx, oy + ty, oz + tz};
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = p[a] > d.S[a] - 1 ? d.S[a] - 1 : p[a];
    gc = ((long)p[0] * d.S[1] + p[1]) * d.S[2] + p[2];
  }

  __device__ __forceinline__ const double* cur() const { return lds + buf * G::LVOL; }

  __device__ __forceinline__ void stage(int f) {
    buf ^= 1;
    double* L = lds + buf * G::LVOL;
    const double* u = F + (long)f * P;
    double v[G::PER];
#pragma unroll
    for (int k = 0; k < G::PER; ++k) v[k] = u[off[k]];
#pragma unroll
    for (int k = 0; k < G::PER; ++k) L[(int)threadIdx.x + k * G::NT] = v[k];
    __syncthreads();
  }

  __device__ __forceinline__ double center(int f) const { return F[(long)f * P + gc]; }
  __device__ __forceinline__ double val() const { return pin(cur()[c]); }

  __device__ __forceinline__ static constexpr int stride(int a) {
    return a == 0 ? G::SXL : (a == 1 ? G::SYL : 1);
  }

  __device__ __forceinline__ double d1(int a) const {
    const double* L = cur() + c;
    const int s = stride(a);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < G::NTAP; ++k)
      if (k != G::R) acc += w.w1[a][k] * L[(k - G::R) * s];
    return pin(acc);
  }

  __device__ __forceinline__ double d2(int i, int j) const {
    const double* L = cur() + c;
    if (i == j) {
      const int s = stride(i);
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < G::NTAP; ++k) acc += w.w2[i][k] * L[(k - G::R) * s];
      return pin(acc);
    }
    const int si = stride(i), sj = stride(j);
    double acc = 0.0;
#pragma unroll
    for (int bb = 0; bb < G::NTAP; ++bb) {
      if (bb == G::R) continue;
      double in = 0.0;
#pragma unroll
      for (int a = 0; a < G::NTAP; ++a)
        if (a != G::R) in += w.w1[i][a] * L[(a - G::R) * si + (bb - G::R) * sj];
      acc += w.w1[j][bb] * in;
    }
    return pin(acc);
  }

  __device__ __forceinline__ double ko() const {
    const double* L = cur() + c;
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int s = stride(a);
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < G::NKO; ++k) t += w.wko[a][k] * L[(k - G::RK) * s];
      acc += t;
    }
    return pin(acc);
  }
};

// Workgroup -> brick, XCD-aware: the dispatcher deals workgroups round-robin over the 8 XCDs,
// so consecutive ids are spread; remap so that each XCD receives a contiguous run of bricks
// (z fastest, then y, then x) and halo re-reads hit that XCD's L2.
__device__ __forceinline__ bool brick_of_block(int nbx, int nby, int nbz, int* bx, int* by,
                                               int* bz) {
  const int G = nbx * nby * nbz;
  const int G8 = (G + 7) & ~7;
  const int w = (int)blockIdx.x;
  const int L = (w & 7) * (G8 >> 3) + (w >> 3);
  if (L >= G) return false;
  *bz = L % nbz;
  const int t = L / nbz;
  *by = t % nby;
  *bx = t / nby;
  return true;
}

}  // namespace bssn
