// Block hierarchy transfers: node-centred 2:1 prolongation and the refinement indicator.
//
// Prolongation (3D/src/multipatch/amr/kernels.py:53-108): the 1D refinement keeps coarse node k
// at fine index 2k and puts at 2k+1 the 6-point Lagrange midpoint
//     fine[2k+1] = sum_{j=-2..3} w_j coarse[k + j],  w = (3, -25, 150, 150, -25, 3) / 256
// (the reference forms w with its Lagrange product in float64, which gives these dyadic values
// exactly), with the coarse reads edge-clamped to the parent's padded block; 3D is the tensor
// product applied x, then y, then z. One thread evaluates one destination point for all 24
// fields directly from the parent: per field the x sums are formed first, then the y sum over
// them, then the z sum, each sum left to right from 0.0 and with contraction off -- the
// rounding of the reference's three sequential passes. Even axes are copies.
//
// Indicator (kernels.py:143-158): the largest undivided difference |u(p + e_a) - u(p)| over the
// interior plus one ring, per (block, field), NaN-propagating like a max over an array.
#include <hip/hip_runtime.h>

#include "bssn_hip.h"
#include "bssn_kernels.h"

namespace bssn {

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct Taps {
  int n;          // 1 (coincident node) or 6 (midpoint)
  int idx[6];     // parent node indices along the axis, clamped
};

__device__ __forceinline__ Taps taps(int fine, int S) {
  Taps t;
  const int k = fine >> 1;
  if ((fine & 1) == 0) {
    t.n = 1;
    t.idx[0] = clampi(k, 0, S - 1);
  } else {
    t.n = 6;
#pragma unroll
    for (int j = 0; j < 6; ++j) t.idx[j] = clampi(k - 2 + j, 0, S - 1);
  }
  return t;
}

__constant__ double kWMid[6] = {3.0 / 256.0, -25.0 / 256.0, 150.0 / 256.0,
                                150.0 / 256.0, -25.0 / 256.0, 3.0 / 256.0};

}  // namespace

__global__ __launch_bounds__(256) void k_prolong(double* __restrict__ pool,
                                                 const bssn_prolong_point* __restrict__ pts,
                                                 long n) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const bssn_prolong_point m = pts[e];
  const Taps tx = taps(m.fine[0], m.pS[0]);
  const Taps ty = taps(m.fine[1], m.pS[1]);
  const Taps tz = taps(m.fine[2], m.pS[2]);
  const long sy = m.pS[2], sx = (long)m.pS[1] * m.pS[2];
  const long P = sx * m.pS[0];
  for (int f = 0; f < BSSN_NF; ++f) {
    const double* c = pool + m.parent + (long)f * P;
    double vz = 0.0;
    for (int kz = 0; kz < tz.n; ++kz) {
      double vy = 0.0;
      for (int ky = 0; ky < ty.n; ++ky) {
        const double* row = c + (long)ty.idx[ky] * sy + tz.idx[kz];
        double vx;
        if (tx.n == 1) {
          vx = row[(long)tx.idx[0] * sx];
        } else {
          vx = 0.0;
#pragma unroll
          for (int kx = 0; kx < 6; ++kx) vx = vx + kWMid[kx] * row[(long)tx.idx[kx] * sx];
        }
        vy = ty.n == 1 ? vx : vy + kWMid[ky] * vx;
      }
      vz = tz.n == 1 ? vy : vz + kWMid[kz] * vy;
    }
    pool[m.dst + (long)f * m.dst_stride] = vz;
  }
}

hipError_t launch_prolong(double* pool, const void* pts, long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_prolong, dim3((unsigned)blocks), dim3(256), 0, st, pool,
                     (const bssn_prolong_point*)pts, n);
  return hipGetLastError();
}

namespace {

__device__ __forceinline__ double nanmax(double m, double v) {
  return (v > m || v != v) ? v : m;     // a NaN, once in, stays
}

}  // namespace

// one workgroup of 256 per (block, field): blockIdx.x = 24 b + f
__global__ __launch_bounds__(256) void k_indicator(const double* __restrict__ pool,
                                                   const BlockDesc* __restrict__ blks,
                                                   double* __restrict__ out) {
  const int b = (int)blockIdx.x / BSSN_NF, f = (int)blockIdx.x % BSSN_NF;
  const BlockDesc blk = blks[b];
  const int lo = blk.ng - 1;
  const int C[3] = {blk.S[0] - 2 * blk.ng + 2, blk.S[1] - 2 * blk.ng + 2,
                    blk.S[2] - 2 * blk.ng + 2};
  const long sy = blk.S[2], sx = (long)blk.S[1] * blk.S[2];
  const double* u = pool + blk.off + (long)f * sx * blk.S[0];
  const long total = (long)C[0] * C[1] * C[2];
  double m = 0.0;
  for (long i = threadIdx.x; i < total; i += blockDim.x) {
    const int z = (int)(i % C[2]);
    const int y = (int)((i / C[2]) % C[1]);
    const int x = (int)(i / ((long)C[2] * C[1]));
    const long p = (long)(lo + x) * sx + (long)(lo + y) * sy + (lo + z);
    const double v = u[p];
    if (x + 1 < C[0]) m = nanmax(m, fabs(u[p + sx] - v));
    if (y + 1 < C[1]) m = nanmax(m, fabs(u[p + sy] - v));
    if (z + 1 < C[2]) m = nanmax(m, fabs(u[p + 1] - v));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
  __shared__ double red[4];
  const int wave = (int)threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < (int)blockDim.x / 64; ++w) r = nanmax(r, red[w]);
    out[blockIdx.x] = r;
  }
}

hipError_t launch_indicator(const double* pool, const BlockDesc* blk, int nblk, double* out,
                            hipStream_t st) {
  if (nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_indicator, dim3((unsigned)(nblk * BSSN_NF)), dim3(256), 0, st, pool, blk,
                     out);
  return hipGetLastError();
}

}  // namespace bssn
