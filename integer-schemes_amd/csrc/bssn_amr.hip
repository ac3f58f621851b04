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

#include "bssn_entries.h"
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

// one field of one destination point, with the tap counts of its parity class fixed at compile
// time (OX: x fine index odd, ...)
template <bool OX, bool OY, bool OZ>
__device__ __forceinline__ void prolong_field(double* __restrict__ pool,
                                              const bssn_prolong_point& m, int f) {
#pragma clang fp contract(off)
  constexpr int NY = OY ? 6 : 1, NZ = OZ ? 6 : 1;
  const Taps tx = taps(m.fine[0], m.pS[0]);
  const Taps ty = taps(m.fine[1], m.pS[1]);
  const Taps tz = taps(m.fine[2], m.pS[2]);
  const long sy = m.pS[2], sx = (long)m.pS[1] * m.pS[2];
  const long P = sx * m.pS[0];
  const double* c = pool + m.parent + (long)f * P;
  double vz = 0.0;
#pragma unroll
  for (int kz = 0; kz < NZ; ++kz) {
    double vy = 0.0;
#pragma unroll
    for (int ky = 0; ky < NY; ++ky) {
      const double* row = c + (long)ty.idx[ky] * sy + tz.idx[kz];
      double vx;
      if constexpr (OX) {
        vx = 0.0;
#pragma unroll
        for (int kx = 0; kx < 6; ++kx) vx = vx + kWMid[kx] * row[(long)tx.idx[kx] * sx];
      } else {
        vx = row[(long)tx.idx[0] * sx];
      }
      if constexpr (OY) vy = vy + kWMid[ky] * vx; else vy = vx;
    }
    if constexpr (OZ) vz = vz + kWMid[kz] * vy; else vz = vy;
  }
  pool[m.dst + (long)f * m.dst_stride] = vz;
}

// One thread per entry and field (blockIdx.y = field): the up to 6^3 taps of a point are
// independent loads, and 24 threads per point keep many more of them in flight than one thread
// looping over the fields. Entries sorted by parity class keep a wave on one branch. An entry
// that would reach outside the pool (pool_n doubles) is skipped.
__global__ __launch_bounds__(256) void k_prolong(double* __restrict__ pool,
                                                 const bssn_prolong_point* __restrict__ pts,
                                                 long n, long long pool_n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = (int)blockIdx.y;
  if (e >= n) return;
  const bssn_prolong_point m = pts[e];
  if (!prolong_entry_ok(m, pool_n)) return;
  const int cls = (m.fine[0] & 1) << 2 | (m.fine[1] & 1) << 1 | (m.fine[2] & 1);
  switch (cls) {
    case 0: prolong_field<false, false, false>(pool, m, f); break;
    case 1: prolong_field<false, false, true>(pool, m, f); break;
    case 2: prolong_field<false, true, false>(pool, m, f); break;
    case 3: prolong_field<false, true, true>(pool, m, f); break;
    case 4: prolong_field<true, false, false>(pool, m, f); break;
    case 5: prolong_field<true, false, true>(pool, m, f); break;
    case 6: prolong_field<true, true, false>(pool, m, f); break;
    default: prolong_field<true, true, true>(pool, m, f); break;
  }
}

// One 4x4x4 fine tile and PT_FIELDS consecutive fields per 64-lane workgroup (blockIdx.y =
// field group). Per field, the 7^3 coarse window of the tile (clamped into the parent, as the
// point kernel's taps) is written to LDS, then the x pass X[fx][cy][cz], the y pass
// Y[fx][fy][cz] and, at the tile's points only, the z pass -- each sum over its taps in the
// point kernel's order, so every point gets the value of its bssn_prolong_point entry bit for
// bit. The window's parent offsets and the lane's tile point are formed once per tile; the next
// field's window is loaded into registers while the current field's passes run. A tile or point
// that would reach outside the pool or the point array is skipped (the tile test is uniform
// over the group). Four fields per group: on the 640-block hierarchy's fill, 2 / 3 / 4 / 6 / 8 /
// 24 fields per group take 1.134 / 1.124 / 1.121 / 1.124 / 1.133 / 1.179 ms per fill, one field
// per group 1.352 ms.
#ifndef BSSN_PT_FIELDS
#define BSSN_PT_FIELDS 4
#endif
constexpr int PT_FIELDS = BSSN_PT_FIELDS;
static_assert(BSSN_NF % PT_FIELDS == 0, "field groups must tile the 24 fields");

__global__ __launch_bounds__(64) void k_prolong_tile(double* __restrict__ pool,
                                                     const bssn_prolong_tile* __restrict__ tiles,
                                                     long ntiles,
                                                     const bssn_prolong_tile_point* __restrict__ tp,
                                                     long long ntp, long long pool_n) {
#pragma clang fp contract(off)
  constexpr int NW = 343, NS = (NW + 63) / 64;     // window cells; per-lane slots
  __shared__ double W[NW];
  __shared__ double X[4 * 49];
  __shared__ double Y[16 * 7];
  // XCD-aware: the dispatcher deals workgroups round-robin over the 8 XCDs (gridDim.x is a
  // multiple of 8, so x % 8 picks the XCD); each XCD takes a contiguous run of tiles, whose
  // coarse windows (tiles are sorted by parent and position) share cache lines
  const long G8 = (long)gridDim.x, w = (long)blockIdx.x;
  const long ti = (w & 7) * (G8 >> 3) + (w >> 3);
  if (ti >= ntiles) return;
  const bssn_prolong_tile T = tiles[ti];
  const int lane = (int)threadIdx.x;
  if (!prolong_tile_ok(T, pool_n) || T.first + T.count > ntp) return;
  const int f0 = (int)blockIdx.y * PT_FIELDS;
  const long sy = T.pS[2], sx = (long)T.pS[1] * T.pS[2];
  const long P = sx * T.pS[0];
  const int c0[3] = {(T.origin[0] >> 1) - 2, (T.origin[1] >> 1) - 2, (T.origin[2] >> 1) - 2};
  long woff[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int i = lane + 64 * s;
    const int x = clampi(c0[0] + i / 49, 0, T.pS[0] - 1);
    const int y = clampi(c0[1] + (i / 7) % 7, 0, T.pS[1] - 1);
    const int z = clampi(c0[2] + i % 7, 0, T.pS[2] - 1);
    woff[s] = x * sx + y * sy + z;
  }
  // the lane's point of the tile's first 64 (a tile holds each of its 64 fine points once)
  bssn_prolong_tile_point q0{};
  bool ok0 = false;
  if (lane < T.count) {
    q0 = tp[T.first + lane];
    ok0 = prolong_tile_point_ok(q0, pool_n);
  }
  double wv[NS];
  {
    const double* c = pool + T.parent + (long)f0 * P;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (lane + 64 * s < NW) wv[s] = c[woff[s]];
  }
#pragma unroll 1
  for (int f = f0; f < f0 + PT_FIELDS; ++f) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (lane + 64 * s < NW) W[lane + 64 * s] = wv[s];
    __syncthreads();
    if (f + 1 < f0 + PT_FIELDS) {                     // next field's window, in flight
      const double* c = pool + T.parent + (long)(f + 1) * P;
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (lane + 64 * s < NW) wv[s] = c[woff[s]];
    }
    for (int i = lane; i < 4 * 49; i += 64) {        // x pass
      const int fx = i / 49, jk = i % 49;
      const int fine = T.origin[0] + fx, kx = (fine >> 1) - c0[0];
      double v;
      if (fine & 1) {
        v = 0.0;
#pragma unroll
        for (int t = 0; t < 6; ++t) v = v + kWMid[t] * W[(kx - 2 + t) * 49 + jk];
      } else {
        v = W[kx * 49 + jk];
      }
      X[i] = v;
    }
    __syncthreads();
    for (int i = lane; i < 16 * 7; i += 64) {        // y pass
      const int fx = i / 28, fy = (i / 7) % 4, k = i % 7;
      const int fine = T.origin[1] + fy, ky = (fine >> 1) - c0[1];
      double v;
      if (fine & 1) {
        v = 0.0;
#pragma unroll
        for (int t = 0; t < 6; ++t) v = v + kWMid[t] * X[fx * 49 + (ky - 2 + t) * 7 + k];
      } else {
        v = X[fx * 49 + ky * 7 + k];
      }
      Y[i] = v;
    }
    __syncthreads();
    auto zpass = [&](const bssn_prolong_tile_point& q) {
      const int fz = q.local & 3;
      const double* yr = Y + (q.local >> 2) * 7;
      const int fine = T.origin[2] + fz, kz = (fine >> 1) - c0[2];
      double v;
      if (fine & 1) {
        v = 0.0;
#pragma unroll
        for (int t = 0; t < 6; ++t) v = v + kWMid[t] * yr[kz - 2 + t];
      } else {
        v = yr[kz];
      }
      pool[q.dst + (long)f * q.dst_stride] = v;
    };
    if (ok0) zpass(q0);
    for (long long e = T.first + 64 + lane; e < T.first + T.count; e += 64) {
      const bssn_prolong_tile_point q = tp[e];
      if (prolong_tile_point_ok(q, pool_n)) zpass(q);
    }
  }
}

hipError_t launch_prolong_tiles(double* pool, long long pool_n, const void* tiles, long n,
                                const void* tp, long long ntp, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long g = (n + 7) & ~7L;
  if (g > 0x7fffffffL) return hipErrorInvalidValue;       // beyond one grid dimension
  hipLaunchKernelGGL(k_prolong_tile, dim3((unsigned)g, BSSN_NF / PT_FIELDS), dim3(64), 0, st,
                     pool, (const bssn_prolong_tile*)tiles, n, (const bssn_prolong_tile_point*)tp,
                     ntp, pool_n);
  return hipGetLastError();
}

hipError_t launch_prolong(double* pool, long long pool_n, const void* pts, long n,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_prolong, dim3((unsigned)blocks, BSSN_NF), dim3(256), 0, st, pool,
                     (const bssn_prolong_point*)pts, n, pool_n);
  return hipGetLastError();
}

// Separable whole-block prolongation: one workgroup per (job, field, x slab), a slab being
// PB_SLABS-th of the padded block's fine x planes. The x pass fills X[rx][ky][kz] (the slab's
// fine x planes, coarse y and z window) from the parent, the y pass Y[rx][ry][kz] from X, and
// the z sums are formed only at the points the job writes. Every fine x plane depends only on
// its own coarse taps, so slabs share no work. Window index k stands for parent node kmin + k,
// read clamped, so the taps of every sum are the clamped nodes of the per-point kernel, summed
// in the same order. A job that is not a valid block of at most PB_WMAX padded nodes per axis
// inside the pool (pool_n doubles) is skipped (the test is uniform across the workgroup, so no
// thread reaches a barrier alone). Eight slabs of 256 threads (19.6 KB LDS, several workgroups
// per CU): on the 384-block hierarchy (16^3 blocks) one fill takes 0.801 ms, against 0.801 /
// 0.811 ms with 2 slabs of 512 / 4 of 256, and 0.883 ms with one 1024-thread workgroup per
// (job, field) holding the whole block's passes (157 KB LDS, one workgroup per CU).
#ifndef BSSN_PB_SLABS
#define BSSN_PB_SLABS 8
#endif
#ifndef BSSN_PB_THREADS
#define BSSN_PB_THREADS 256
#endif
constexpr int PB_SLABS = BSSN_PB_SLABS;
constexpr int PB_SW = (PB_WMAX + PB_SLABS - 1) / PB_SLABS;   // fine x planes per slab, at most

__global__ __launch_bounds__(BSSN_PB_THREADS) void k_prolong_block(
    double* __restrict__ pool, const bssn_prolong_block* __restrict__ jobs, long long pool_n) {
#pragma clang fp contract(off)
  __shared__ double X[PB_SW * PB_KMAX * PB_KMAX];
  __shared__ double Y[PB_SW * PB_WMAX * PB_KMAX];
  const int slab = (int)(blockIdx.x % PB_SLABS);
  const long jf = blockIdx.x / PB_SLABS;
  const bssn_prolong_block jb = jobs[jf / BSSN_NF];
  if (!block_job_ok(jb, pool_n)) return;
  const int f = (int)(jf % BSSN_NF);
  const int W = jb.BS + 2 * jb.ng;
  const int sw = (W + PB_SLABS - 1) / PB_SLABS;
  const int x0 = slab * sw, nx = min(W, x0 + sw) - x0;
  if (nx <= 0) return;
  const long sy = W, sx = (long)W * W, P = sx * W;
  int r0[3], kmin[3], K[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r0[a] = jb.ng + jb.corner[a] * jb.BS;
    kmin[a] = (r0[a] >> 1) - 2;
    K[a] = ((r0[a] + W - 1) >> 1) + 3 - kmin[a] + 1;
  }
  const double* c = pool + jb.parent + (long)f * P;
  const int nX = nx * K[1] * K[2];
  for (int i = (int)threadIdx.x; i < nX; i += (int)blockDim.x) {
    const int kz = i % K[2], ky = (i / K[2]) % K[1], xl = i / (K[2] * K[1]);
    const int rx = r0[0] + x0 + xl;
    const int ny = clampi(kmin[1] + ky, 0, W - 1), nz = clampi(kmin[2] + kz, 0, W - 1);
    const double* row = c + (long)ny * sy + nz;
    double v;
    if ((rx & 1) == 0) {
      v = row[(long)clampi(rx >> 1, 0, W - 1) * sx];
    } else {
      v = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) v = v + kWMid[j] * row[(long)clampi((rx >> 1) - 2 + j, 0, W - 1) * sx];
    }
    X[(xl * PB_KMAX + ky) * PB_KMAX + kz] = v;
  }
  __syncthreads();
  const int nY = nx * W * K[2];
  for (int i = (int)threadIdx.x; i < nY; i += (int)blockDim.x) {
    const int kz = i % K[2], y = (i / K[2]) % W, xl = i / (K[2] * W);
    const int ry = r0[1] + y;
    const double* col = X + (long)xl * PB_KMAX * PB_KMAX + kz;
    double v;
    if ((ry & 1) == 0) {
      v = col[((ry >> 1) - kmin[1]) * PB_KMAX];
    } else {
      v = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) v = v + kWMid[j] * col[((ry >> 1) - 2 + j - kmin[1]) * PB_KMAX];
    }
    Y[(xl * PB_WMAX + y) * PB_KMAX + kz] = v;
  }
  __syncthreads();
  const int lo = jb.ng, hi = jb.ng + jb.BS;
  double* out = pool + jb.dst + (long)f * P;
  for (int i = (int)threadIdx.x; i < nx * W * W; i += (int)blockDim.x) {
    const int z = i % W, y = (i / W) % W, xl = i / (W * W), x = x0 + xl;
    const bool ix = x >= lo && x < hi, iy = y >= lo && y < hi, iz = z >= lo && z < hi;
    if (ix && iy && iz) {
      if (!(jb.skip & 64)) continue;
    } else {
      // a face slab: ghost along exactly one axis, interior along the other two
      if (iy && iz && (jb.skip & (x < lo ? 1 : 2))) continue;
      if (ix && iz && (jb.skip & (y < lo ? 4 : 8))) continue;
      if (ix && iy && (jb.skip & (z < lo ? 16 : 32))) continue;
    }
    const int rz = r0[2] + z;
    const double* yy = Y + (xl * PB_WMAX + y) * PB_KMAX;
    double v;
    if ((rz & 1) == 0) {
      v = yy[(rz >> 1) - kmin[2]];
    } else {
      v = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) v = v + kWMid[j] * yy[(rz >> 1) - 2 + j - kmin[2]];
    }
    out[((long)x * W + y) * W + z] = v;
  }
}

hipError_t launch_prolong_blocks(double* pool, long long pool_n, const void* jobs, long n,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n > 0x7fffffffL / (BSSN_NF * PB_SLABS)) return hipErrorInvalidValue;   // grid too large
  hipLaunchKernelGGL(k_prolong_block, dim3((unsigned)(n * BSSN_NF * PB_SLABS)),
                     dim3(BSSN_PB_THREADS), 0, st, pool, (const bssn_prolong_block*)jobs, pool_n);
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
