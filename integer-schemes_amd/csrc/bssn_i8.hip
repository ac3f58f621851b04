// INT8-split derivative bundle (Ozaki scheme I on the INT8 matrix cores) for FD order 4.
//
// Precision form of the derivative bundle (3D/src/bssn3d/derivative_bundle.py:31-48): the
// 138 derivatives are taken from 2D slabs of one field, and on each slab the stencil is applied
// as a banded integer matrix by INT8 MFMA with INT32 accumulation, on int8 digits of the values.
// The reference's precision form (Ozaki scheme II with residues and CRT, for its Maxwell-Chern-
// Simons solver only: 2D/src/mcs2d/schemes/ozaki.py:257-562, 3D/src/mcs3d/schemes/ozaki.py:
// 162-282) contributes the per-block power-of-two scale (ozaki.py:316-323) and the banded
// stencil matrix (ozaki.py:286-306); the digits here are an Ozaki-I split instead of residues.
//
// One wave computes a 28 x 28 tile of outputs of one plane of one field from the 32 x 32 box
// around it (coordinates clamped to the grid, the reference's edge padding):
//   box values v -> c = 0.5 max + 0.5 min (stencils annihilate constants), sigma = 2^(6-E) with
//   max|v - c| = f 2^E, r0 = (v - c) sigma, q_s = rint(r_s), r_{s+1} = 128 (r_s - q_s):
//   |q_s| <= 64, every step exact; (v - c) sigma = sum_s q_s 2^(-7 s) + O(2^(-7 k)).
//   Per digit plane q_s (int8, 32 x 32):
//     B q   (rows: outputs along axis 1)   = v_mfma_i32_32x32x32_i8(band, q)
//     q B^T (columns: outputs along axis 2) = v_mfma_i32_32x32x32_i8(q^T via LDS, band)
//     B q B^T (mixed): the INT32 B q is split into two int8 digits (|B q| <= 18 * 64) and each
//           is multiplied by B^T, so the mixed contraction is exact too.
//   Recombination in FP64: acc += P_s 2^(-7 s) (exact terms), D = (acc 2^-e) * scale.
// The integer stencils are those of common/src/mcs_common/derivatives.py:33-49 times 12
// (first and second derivatives) and 144 (mixed).
//
// Plane orientations: (y, z) planes at each x give d_y, d_yy, d_z, d_zz, d_yz; (x, z) planes at
// each y give d_x, d_xx, d_xz; (x, y) planes at each z give d_xy.
#include <hip/hip_runtime.h>

#include <math.h>

#include "bssn_kernels.h"
#include "bssn_params.h"

namespace bssn {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int I8_TILE = 28, I8_BOX = 32;

// fields with second derivatives (alpha, chi, gt0..5, beta0..2) and the others
__constant__ int kTwoFields[11] = {ALPHA, CHI, GT0, GT1, GT2, GT3, GT4, GT5, BETA0, BETA1, BETA2};
__constant__ int kOneFields[13] = {KTR, AT0, AT1, AT2, AT3, AT4, AT5, GTI0, GTI0 + 1, GTI0 + 2,
                                   BB0, BB1, BB2};

// taps at offsets -2..2 (index t = 0..4)
__device__ __forceinline__ int tap1(int t) {
  return t == 0 ? 1 : (t == 1 ? -8 : (t == 3 ? 8 : (t == 4 ? -1 : 0)));
}
__device__ __forceinline__ int tap2(int t) {
  return (t == 0 || t == 4) ? -1 : ((t == 1 || t == 3) ? 16 : -30);
}

// The lane's 16 bytes of the banded matrix: row r = lane % 32 (an output), positions
// k = 16 (lane / 32) + j. The same bytes serve as the A operand B (rows r) and as the
// B operand B^T (columns r).
template <int W>
__device__ __forceinline__ v4i band_operand(int lane) {
  const int r = lane & 31, k0 = 16 * (lane >> 5);
  unsigned w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int t = k0 + j - r;
    int v = 0;
    if (r < I8_TILE && t >= 0 && t <= 4) v = W == 1 ? tap1(t) : tap2(t);
    w[j >> 2] |= ((unsigned)v & 0xffu) << (8 * (j & 3));
  }
  return v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

__device__ __forceinline__ v4i pack16(const int* q) {
  unsigned w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int j = 0; j < 16; ++j) w[j >> 2] |= ((unsigned)q[j] & 0xffu) << (8 * (j & 3));
  return v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x = fmax(x, __shfl_xor(x, m, 64));
  return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x = fmin(x, __shfl_xor(x, m, 64));
  return x;
}

// C-layout of the 32 x 32 result: lane holds column lane % 32, rows row_of(i, lane)
__device__ __forceinline__ int row_of(int i, int lane) {
  return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
}

struct I8Args {
  const double* F;
  double* D;
  int S[3];
  long P;
  double s1[3], s2[3], sm[3];
  int nslice;
};

// Launch geometry of one (ORIENT, TWO) pass over a grid of padded extents S: the number of
// wave tiles (fields x tiles along the plane's two axes x planes).
template <int ORIENT, bool TWO>
__host__ __device__ inline long i8_tiles(const int* S) {
  const int ax1 = ORIENT == 0 ? 1 : 0, ax2 = ORIENT == 2 ? 1 : 2;
  const int axp = ORIENT == 0 ? 0 : (ORIENT == 1 ? 1 : 2);
  const long nt1 = (S[ax1] + I8_TILE - 1) / I8_TILE, nt2 = (S[ax2] + I8_TILE - 1) / I8_TILE;
  return (TWO ? 11L : 13L) * nt1 * nt2 * S[axp];
}

// ORIENT 0: (y, z) planes at each x; 1: (x, z) planes at each y; 2: (x, y) planes at each z.
// TWO: the fields with second derivatives (else first derivatives only).
// One wave: tile w (plane fastest, then tile along axis 2, along axis 1, field) of the grid a.
template <int ORIENT, bool TWO>
__device__ __forceinline__ void i8_tile(const I8Args& a, long w, int lane, signed char* tq,
                                        signed char* th, signed char* tl) {
  constexpr int AX1 = ORIENT == 0 ? 1 : 0;
  constexpr int AX2 = ORIENT == 2 ? 1 : 2;
  constexpr int AXP = ORIENT == 0 ? 0 : (ORIENT == 1 ? 1 : 2);
  constexpr bool D1A = ORIENT != 2;                 // first derivative along axis 1
  constexpr bool D2A = TWO && ORIENT != 2;          // second derivative along axis 1
  constexpr bool D1B = ORIENT == 0;                 // along axis 2 (z, from the (y, z) planes)
  constexpr bool D2B = TWO && ORIENT == 0;
  constexpr bool MIX = TWO;
  constexpr int NF = TWO ? 11 : 13;

  const int n1 = a.S[AX1], n2 = a.S[AX2], npl = a.S[AXP];
  const long st[3] = {(long)a.S[1] * a.S[2], (long)a.S[2], 1L};
  const int nt1 = (n1 + I8_TILE - 1) / I8_TILE, nt2 = (n2 + I8_TILE - 1) / I8_TILE;
  const int pl = (int)(w % npl);
  w /= npl;
  const int t2 = (int)(w % nt2);
  w /= nt2;
  const int t1 = (int)(w % nt1);
  const int fi = (int)(w / nt1);
  if (fi >= NF) return;
  const int f = TWO ? kTwoFields[fi] : kOneFields[fi];
  const int o1 = t1 * I8_TILE, o2 = t2 * I8_TILE;
  const double* u = a.F + (long)f * a.P + (long)pl * st[AXP];

  // box values in the B-operand distribution: column c = lane % 32, rows 16 (lane / 32) + j
  const int c = lane & 31, hh = lane >> 5;
  int i2 = o2 - 2 + c;
  i2 = i2 < 0 ? 0 : (i2 > n2 - 1 ? n2 - 1 : i2);
  double r[16];
  double mx = -INFINITY, mn = INFINITY;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    int i1 = o1 - 2 + 16 * hh + j;
    i1 = i1 < 0 ? 0 : (i1 > n1 - 1 ? n1 - 1 : i1);
    r[j] = u[(long)i1 * st[AX1] + (long)i2 * st[AX2]];
    mx = fmax(mx, r[j]);
    mn = fmin(mn, r[j]);
  }
  mx = wave_max(mx);
  mn = wave_min(mn);
  const double off = 0.5 * mx + 0.5 * mn;
  double m = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    r[j] = r[j] - off;
    m = fmax(m, fabs(r[j]));
  }
  m = wave_max(m);
  int e = 0;
  if (m != 0.0) {
    int E;
    (void)frexp(m, &E);
    e = 6 - E;
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) r[j] = ldexp(r[j], e);

  const v4i band1 = band_operand<1>(lane);
  const v4i band2 = band_operand<2>(lane);
  double acc1a[16], acc2a[16], acc1b[16], acc2b[16], accm[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc1a[i] = acc2a[i] = acc1b[i] = acc2b[i] = accm[i] = 0.0;

  for (int s = 0; s < a.nslice; ++s) {
    int q[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const double d = rint(r[j]);
      q[j] = (int)d;
      r[j] = (r[j] - d) * 128.0;
    }
    const double ws = ldexp(1.0, -7 * s);
    const v4i qb = pack16(q);
    const v16i z16 = {};
    v16i p1 = z16;
    if (D1A || MIX) p1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(band1, qb, z16, 0, 0, 0);
    if (D2A) {
      const v16i p2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(band2, qb, z16, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 16; ++i) acc2a[i] = acc2a[i] + (double)p2[i] * ws;
    }
    if (D1A) {
#pragma unroll
      for (int i = 0; i < 16; ++i) acc1a[i] = acc1a[i] + (double)p1[i] * ws;
    }
    if (D1B || D2B) {
      // q^T in the A distribution: row = lane % 32 (axis 1), positions along axis 2
#pragma unroll
      for (int j = 0; j < 16; ++j) tq[(16 * hh + j) * I8_BOX + c] = (signed char)q[j];
      __builtin_amdgcn_wave_barrier();
      const v4i qa = *(const v4i*)&tq[c * I8_BOX + 16 * hh];
      __builtin_amdgcn_wave_barrier();
      if (D1B) {
        const v16i p = __builtin_amdgcn_mfma_i32_32x32x32_i8(qa, band1, z16, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc1b[i] = acc1b[i] + (double)p[i] * ws;
      }
      if (D2B) {
        const v16i p = __builtin_amdgcn_mfma_i32_32x32x32_i8(qa, band2, z16, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) acc2b[i] = acc2b[i] + (double)p[i] * ws;
      }
    }
    if (MIX) {
      // B q (INT32, |.| <= 18 * 64) as two int8 digits hi * 128 + lo, each times B^T
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int v = p1[i];
        const int lo = ((v + 64) & 127) - 64;
        const int hi = (v - lo) >> 7;
        const int rr = row_of(i, lane);
        th[rr * I8_BOX + c] = (signed char)hi;
        tl[rr * I8_BOX + c] = (signed char)lo;
      }
      __builtin_amdgcn_wave_barrier();
      const v4i ah = *(const v4i*)&th[c * I8_BOX + 16 * hh];
      const v4i al = *(const v4i*)&tl[c * I8_BOX + 16 * hh];
      __builtin_amdgcn_wave_barrier();
      const v16i mh = __builtin_amdgcn_mfma_i32_32x32x32_i8(ah, band1, z16, 0, 0, 0);
      const v16i ml = __builtin_amdgcn_mfma_i32_32x32x32_i8(al, band1, z16, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 16; ++i) accm[i] = accm[i] + (double)(mh[i] * 128 + ml[i]) * ws;
    }
  }

  // outputs: rows along axis 1, columns along axis 2 (lane's column c, rows row_of(i))
  double* Dp = a.D;
  const long plo = (long)pl * st[AXP];
  auto put = [&](int d, const double* acc, double scale, bool rows_out, bool cols_out) {
    double* dst = Dp + (long)d * a.P + plo;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int rr = row_of(i, lane);
      const int g1 = rows_out ? o1 + rr : o1 + rr - 2;
      const int g2 = cols_out ? o2 + c : o2 + c - 2;
      const bool ok1 = rows_out ? rr < I8_TILE : (rr >= 2 && rr < I8_TILE + 2);
      const bool ok2 = cols_out ? c < I8_TILE : (c >= 2 && c < I8_TILE + 2);
      if (ok1 && ok2 && g1 < n1 && g2 < n2)
        dst[(long)g1 * st[AX1] + (long)g2 * st[AX2]] = ldexp(acc[i], -e) * scale;
    }
  };
  const int slot = TWO ? fi : -1;
  if (D1A) put(3 * f + AX1, acc1a, a.s1[AX1], true, false);
  if (D2A) put(72 + 6 * slot + (AX1 == 0 ? 0 : 3), acc2a, a.s2[AX1], true, false);
  if (D1B) put(3 * f + AX2, acc1b, a.s1[AX2], false, true);
  if (D2B) put(72 + 6 * slot + 5, acc2b, a.s2[AX2], false, true);
  if (MIX) {
    // pair index of (AX1, AX2): xy 1, xz 2, yz 4; scale index AX1 + AX2 - 1
    const int pair = AX1 == 0 ? (AX2 == 1 ? 1 : 2) : 4;
    put(72 + 6 * slot + pair, accm, a.sm[AX1 + AX2 - 1], true, true);
  }
}

#define BSSN_I8_LDS                                                          \
  __shared__ __attribute__((aligned(16))) signed char tq[I8_BOX * I8_BOX];   \
  __shared__ __attribute__((aligned(16))) signed char th[I8_BOX * I8_BOX];   \
  __shared__ __attribute__((aligned(16))) signed char tl[I8_BOX * I8_BOX]

// Workgroup -> tile, XCD-aware: the dispatcher deals workgroups round-robin over the 8 XCDs;
// remap so that each XCD takes a contiguous run of tiles (consecutive planes), whose loads and
// stores share cache lines along the plane axis (a line of 16 doubles along z holds the same
// point of 16 consecutive (x, y) planes). The grid is a multiple of 8; tiles past the last are
// empty (i8_tile returns on them).
__device__ __forceinline__ long xcd_tile() {
  const long G8 = (long)gridDim.x;
  const long w = (long)blockIdx.x;
  return (w & 7) * (G8 >> 3) + (w >> 3);
}

template <int ORIENT, bool TWO>
__global__ __launch_bounds__(64) void k_i8_plane(I8Args a) {
  BSSN_I8_LDS;
  i8_tile<ORIENT, TWO>(a, xcd_tile(), (int)threadIdx.x, tq, th, tl);
}

// The same pass over every block of a batch: block b owns wave tiles [pre[b], pre[b + 1]) of
// the launch; its bundle is the 138 arrays at D + blk[b].doff and its stencil scales are formed
// from its spacings with the host's rounding, so each block's bundle equals the single-grid
// bundle of that block bit for bit.
template <int ORIENT, bool TWO>
__global__ __launch_bounds__(64) void k_i8_plane_batch(const double* pool, double* D,
                                                       const BlockDesc* __restrict__ blks,
                                                       const long long* __restrict__ pre,
                                                       int nblk, int nslice) {
  BSSN_I8_LDS;
  const long long w = (long long)xcd_tile();
  int lo = 0, hi = nblk;                 // pre[lo] <= w < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= w) lo = mid;
    else hi = mid;
  }
  const int b = __builtin_amdgcn_readfirstlane(lo);
  const BlockDesc blk = blks[b];
  I8Args a;
  a.F = pool + blk.off;
  a.D = D + blk.doff;
  for (int i = 0; i < 3; ++i) a.S[i] = blk.S[i];
  a.P = (long)blk.S[0] * blk.S[1] * blk.S[2];
  const Scales sc = make_scales(blk.h, 4, 0.0);
  for (int i = 0; i < 3; ++i) {
    a.s1[i] = uniform(sc.s1[i]);
    a.s2[i] = uniform(sc.s2[i]);
    a.sm[i] = uniform(sc.sm[i]);
  }
  a.nslice = nslice;
  i8_tile<ORIENT, TWO>(a, (long)(w - pre[b]), (int)threadIdx.x, tq, th, tl);
}

template <int ORIENT, bool TWO>
static hipError_t launch_i8_t(const I8Args& a, hipStream_t st) {
  const long n = (i8_tiles<ORIENT, TWO>(a.S) + 7) & ~7L;
  if (n > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_i8_plane<ORIENT, TWO>), dim3((unsigned)n), dim3(64), 0, st, a);
  return hipGetLastError();
}

// wave tiles of pass k (0: (y,z) planes, fields with second derivatives; 1: (y,z), the others;
// 2, 3: the same on (x,z) planes; 4: (x,y) planes, mixed d_xy) on a grid of padded extents S
long i8_pass_tiles(int k, const int* S) {
  switch (k) {
    case 0: return i8_tiles<0, true>(S);
    case 1: return i8_tiles<0, false>(S);
    case 2: return i8_tiles<1, true>(S);
    case 3: return i8_tiles<1, false>(S);
    default: return i8_tiles<2, true>(S);
  }
}

template <int ORIENT, bool TWO>
static hipError_t launch_i8_batch_t(const double* pool, double* D, const BlockDesc* blk,
                                    const long long* pre, long long n, int nblk, int nslice,
                                    hipStream_t st) {
  if (n == 0) return hipSuccess;
  n = (n + 7) & ~7LL;
  if (n > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_i8_plane_batch<ORIENT, TWO>), dim3((unsigned)n), dim3(64), 0, st, pool,
                     D, blk, pre, nblk, nslice);
  return hipGetLastError();
}

// the five passes over a batch; pre: 5 prefix arrays of nblk + 1 entries (device), total: the
// five tile counts (host)
hipError_t launch_batch_bundle_i8(const double* pool, double* D, const BlockDesc* blk,
                                  const long long* pre, const long long* total, int nblk,
                                  int nslice, hipStream_t st) {
  const long long* p = pre;
  const long long m = nblk + 1;
  hipError_t e;
  if ((e = launch_i8_batch_t<0, true>(pool, D, blk, p, total[0], nblk, nslice, st))) return e;
  if ((e = launch_i8_batch_t<0, false>(pool, D, blk, p + m, total[1], nblk, nslice, st))) return e;
  if ((e = launch_i8_batch_t<1, true>(pool, D, blk, p + 2 * m, total[2], nblk, nslice, st)))
    return e;
  if ((e = launch_i8_batch_t<1, false>(pool, D, blk, p + 3 * m, total[3], nblk, nslice, st)))
    return e;
  return launch_i8_batch_t<2, true>(pool, D, blk, p + 4 * m, total[4], nblk, nslice, st);
}

hipError_t launch_deriv_bundle_i8(const double* F, double* D, const Dims& d, const Scales& w,
                                  int nslice, hipStream_t st) {
  I8Args a;
  a.F = F;
  a.D = D;
  for (int i = 0; i < 3; ++i) {
    a.S[i] = d.S[i];
    a.s1[i] = w.s1[i];
    a.s2[i] = w.s2[i];
    a.sm[i] = w.sm[i];
  }
  a.P = d.P;
  a.nslice = nslice;
  hipError_t e;
  if ((e = launch_i8_t<0, true>(a, st)) != hipSuccess) return e;
  if ((e = launch_i8_t<0, false>(a, st)) != hipSuccess) return e;
  if ((e = launch_i8_t<1, true>(a, st)) != hipSuccess) return e;
  if ((e = launch_i8_t<1, false>(a, st)) != hipSuccess) return e;
  return launch_i8_t<2, true>(a, st);
}

}  // namespace bssn
