// The materialised two-pass RHS path and the auxiliary kernels of the step (gfx950).
//
// Two-pass path (the data flow of the reference's verbatim scheme, rhs.py:102-134):
//   pass 1  k_deriv_bundle : 24 fields -> 138 derivative arrays in HBM
//           (derivative_bundle.py:31-48; device analogue of cuda/deriv_2p5d.cu / deriv_smem.cu)
//   pass 2  k_algebra      : 24 fields + 138 derivatives -> 24 right-hand sides, pointwise
//           (_bssn_rhs_generated.py:bssn_rhs_algebra)
//   KO      k_ko_add       : OUT += Kreiss-Oliger over all 24 fields (evolve.py:76-84)
// Auxiliary:
//   k_periodic_sync (evolve.py:65-73), k_enforce (evolve.py:93-119),
//   k_constraints (constraints.py:38-51) and k_sumsq for the interior norms (:53-62).
#include <hip/hip_runtime.h>

#include <math.h>

#include <mutex>
#include <type_traits>

#include "bssn_algebra.h"
#include "bssn_brick.h"
#include "bssn_kernels.h"
#include "bssn_params.h"

namespace bssn {

// field(f, slot) for f = 0..23 in a loop that is not unrolled: one field per iteration, or with
// two staging slots two per iteration, so that each slot index is a compile-time constant
template <int DEPTH, class Fn>
__device__ __forceinline__ void for_fields(Fn&& field) {
  if constexpr (DEPTH == 1) {
#pragma unroll 1
    for (int f = 0; f < BSSN_NF; ++f) field(f, std::integral_constant<int, 0>());
  } else {
    static_assert(DEPTH == 2 && BSSN_NF % 2 == 0, "fields are stepped in pairs");
#pragma unroll 1
    for (int f = 0; f < BSSN_NF; f += 2) {
      field(f, std::integral_constant<int, 0>());
      field(f + 1, std::integral_constant<int, 1>());
    }
  }
}

// second-derivative slot of a field (alpha, chi, gt0..5, beta0..2), -1 otherwise
__host__ __device__ constexpr int s2slot(int f) {
  return f == ALPHA ? 0 : (f == CHI ? 1 : (f >= GT0 && f <= GT5 ? 2 + f - GT0
                                          : (f >= BETA0 && f <= BETA2 ? 8 + f - BETA0 : -1)));
}

// ------------------------------------------------------------------------- pass 1: bundle
// One kernel body, three stores (so every store sees the same derivative values):
//   BS_F64  the fp64 bundle;
//   BS_Q32  / BS_Q48  block-floating-point storage (common/src/mcs_common/bfp_compress.py:65-90,
//           bfp48.py:7-25): m = clip(rint(d 2^e), -2^(k-1), 2^(k-1) - 1), e per array, as int32
//           (k <= 32) or as three 16-bit two's-complement limbs (k <= 48).
enum { BS_F64 = 0, BS_Q32 = 2, BS_Q48 = 3 };

__device__ __forceinline__ double nanmax_d(double m, double v) {
  return (v > m || v != v) ? v : m;
}

template <int STORE>
__device__ __forceinline__ void bundle_put(const BundleOut& o, int slot, long P, long gc,
                                           double v, bool valid) {
  if constexpr (STORE == BS_F64) {
    if (valid) o.D[(long)slot * P + gc] = v;
  } else {
    if (!valid) return;
    // saturate to [-2^(k-1), 2^(k-1) - 1]; a NaN (which an integer mantissa cannot hold) stores
    // as the negative saturation value, like -inf, so the conversion below is always defined
    double m = rint(ldexp(v, o.e[slot]));
    m = !(m >= -o.lim) ? -o.lim : (m > o.lim - 1.0 ? o.lim - 1.0 : m);
    const long i = (long)slot * P + gc;
    if constexpr (STORE == BS_Q32) {
      ((int*)o.Q)[i] = (int)m;
    } else {
      const long long mi = (long long)m;
      unsigned short* q = (unsigned short*)o.Q + 3 * i;
      q[0] = (unsigned short)(mi & 0xFFFF);
      q[1] = (unsigned short)((mi >> 16) & 0xFFFF);
      q[2] = (unsigned short)((mi >> 32) & 0xFFFF);
    }
  }
}

template <int ORD, int STORE>
__global__ __launch_bounds__(64) void k_deriv_bundle(const double* __restrict__ F, BundleOut o,
                                                     Dims d, Scales w) {
  using G = Geo<ORD, false>;
  __shared__ double lds[G::VOL];
  int ox, oy, oz;
  if (!brick_origin<G>(d, &ox, &oy, &oz)) return;
  const int lane = (int)threadIdx.x;
  const bool valid = lane_valid<G>(d, ox, oy, oz, lane);
  // FD4 fp64 / BFP32 stores: two fields in flight, stepped in pairs so each staging slot is a
  // constant (256^3: two-pass RHS with KO 28.6 -> 26.8 ms, BFP32 with fixed exponents 15.9 ->
  // 13.1 ms). The BFP48 store and FD order 6 keep one (their extra registers cost more waves
  // per SIMD than the second field gains: BFP48 19.0 -> 21.1 ms with two).
  constexpr int DEPTH = ORD == 4 && STORE != BS_Q48 ? 2 : 1;
  WaveSrc<ORD, false, CLAMP, StateOrder, DEPTH> s(lds, F, w, d, ox, oy, oz, lane);
  auto field = [&](int f, auto slot) {
    s.template stage_slot<decltype(slot)::value>(f);
    double v[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = s.d1(a);
    const int sl = s2slot(f);
    if (sl >= 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) v[3 + q] = s.d2(pi_(q), pj_(q));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) bundle_put<STORE>(o, 3 * f + a, d.P, s.gc, v[a], valid);
    if (sl >= 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q)
        bundle_put<STORE>(o, 72 + 6 * sl + q, d.P, s.gc, v[3 + q], valid);
    }
    __builtin_amdgcn_wave_barrier();
  };
  for_fields<DEPTH>(field);
}

template <int STORE>
static void launch_bundle_store(const double* F, const BundleOut& o, const Dims& d,
                                const Scales& w, int ord, hipStream_t st) {
  with_order(ord, [&](auto oc) {
    constexpr int O = decltype(oc)::value;
    hipLaunchKernelGGL((k_deriv_bundle<O, STORE>), dim3(brick_blocks<Geo<O, false>>(d)),
                       dim3(64), 0, st, F, o, d, w);
  });
}

hipError_t launch_deriv_bundle(const double* F, double* D, const Dims& d, const Scales& w,
                               int ord, hipStream_t st) {
  BundleOut o{};
  o.D = D;
  launch_bundle_store<BS_F64>(F, o, d, w, ord, st);
  return hipGetLastError();
}

__global__ void k_clear_u64(unsigned long long* p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0ull;
}

__global__ void k_put_scales(double* dst, BfpScales s) {
  const int i = (int)threadIdx.x;
  if (i < BSSN_ND) dst[i] = s.inv[i];
}

// Per-array maxima with few atomics: each wave walks many bricks, keeps the running maxima in
// registers (lane l holds arrays l, l + 64, l + 128) and ends with one atomicMax per array.
template <int ORD>
__global__ __launch_bounds__(64) void k_bundle_absmax(const double* __restrict__ F,
                                                      unsigned long long* __restrict__ amax,
                                                      Dims d, Scales w) {
  using G = Geo<ORD, false>;
  __shared__ double lds[G::VOL];
  const int lane = (int)threadIdx.x;
  const int nbx = (d.n[0] + G::BX - 1) / G::BX, nby = (d.n[1] + G::BY - 1) / G::BY,
            nbz = (d.n[2] + G::BZ - 1) / G::BZ;
  const int nb = nbx * nby * nbz;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  auto keep = [&](int slot, double v, bool valid) {
    double a = valid ? fabs(v) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = nanmax_d(a, __shfl_xor(a, off, 64));
    if (lane == (slot & 63)) {
      const int q = slot >> 6;
      if (q == 0) r0 = nanmax_d(r0, a);
      else if (q == 1) r1 = nanmax_d(r1, a);
      else r2 = nanmax_d(r2, a);
    }
  };
  for (int L = (int)blockIdx.x; L < nb; L += (int)gridDim.x) {
    const int bz = L % nbz, t = L / nbz, by = t % nby, bx = t / nby;
    const int ox = d.lo[0] + bx * G::BX, oy = d.lo[1] + by * G::BY, oz = d.lo[2] + bz * G::BZ;
    const bool valid = lane_valid<G>(d, ox, oy, oz, lane);
    // one field in flight (two measured no faster here)
    WaveSrc<ORD, false, CLAMP, StateOrder> s(lds, F, w, d, ox, oy, oz, lane);
    auto field = [&](int f, auto slot) {
      s.template stage_slot<decltype(slot)::value>(f);
      double v[9];
#pragma unroll
      for (int a = 0; a < 3; ++a) v[a] = s.d1(a);
      const int sl = s2slot(f);
      if (sl >= 0) {
#pragma unroll
        for (int q = 0; q < 6; ++q) v[3 + q] = s.d2(pi_(q), pj_(q));
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) keep(3 * f + a, v[a], valid);
      if (sl >= 0) {
#pragma unroll
        for (int q = 0; q < 6; ++q) keep(72 + 6 * sl + q, v[3 + q], valid);
      }
      __builtin_amdgcn_wave_barrier();
    };
    for_fields<1>(field);
  }
  atomicMax(amax + lane, (unsigned long long)__double_as_longlong(r0));
  atomicMax(amax + 64 + lane, (unsigned long long)__double_as_longlong(r1));
  if (128 + lane < BSSN_ND)
    atomicMax(amax + 128 + lane, (unsigned long long)__double_as_longlong(r2));
}

hipError_t launch_bundle_absmax(const double* F, unsigned long long* amax, const Dims& d,
                                const Scales& w, int ord, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(256), 0, st, amax, BSSN_ND);
  const unsigned grid = 2048;
  with_order(ord, [&](auto oc) {
    constexpr int O = decltype(oc)::value;
    hipLaunchKernelGGL(k_bundle_absmax<O>, dim3(grid), dim3(64), 0, st, F, amax, d, w);
  });
  return hipGetLastError();
}

hipError_t launch_bundle_bfp(const double* F, void* Q, const BundleOut& o_in, const BfpScales& sc,
                             const Dims& d, const Scales& w, int ord, bool wide,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_put_scales, dim3(1), dim3(256), 0, st, (double*)Q, sc);
  BundleOut o = o_in;
  o.Q = (unsigned char*)Q + BFP_HEADER_BYTES;
  if (wide)
    launch_bundle_store<BS_Q48>(F, o, d, w, ord, st);
  else
    launch_bundle_store<BS_Q32>(F, o, d, w, ord, st);
  return hipGetLastError();
}

// ------------------------------------------------------------------------- pass 2: algebra
struct BundleSrc {
  const double* F;
  const double* D;
  long n, p;
  int cur;
  __device__ __forceinline__ void stage(int f) { cur = f; }
  __device__ __forceinline__ double center(int f) const { return F[(long)f * n + p]; }
  __device__ __forceinline__ double val() const { return F[(long)cur * n + p]; }
  __device__ __forceinline__ double d1(int a) const { return D[(long)(3 * cur + a) * n + p]; }
  __device__ __forceinline__ double d2(int i, int j) const {
    return D[(long)(72 + 6 * s2slot(cur) + sym(i, j)) * n + p];
  }
  __device__ __forceinline__ double ko() const { return 0.0; }
};

struct PlainOut {
  double* OUT;
  long n, p;
  __device__ __forceinline__ void put(int f, double v) { OUT[(long)f * n + p] = v; }
  __device__ __forceinline__ void metric_done() {}
};

// the bundle read back from BFP storage: d = m 2^-e, exact (the header holds 2^-e per array)
template <bool WIDE>
struct BundleSrcQ {
  const double* F;
  const double* inv;
  const void* Q;
  long n, p;
  int cur;
  __device__ __forceinline__ void stage(int f) { cur = f; }
  __device__ __forceinline__ double center(int f) const { return F[(long)f * n + p]; }
  __device__ __forceinline__ double val() const { return F[(long)cur * n + p]; }
  __device__ __forceinline__ double get(int slot) const {
    const long i = (long)slot * n + p;
    double m;
    if constexpr (WIDE) {
      const unsigned short* q = (const unsigned short*)Q + 3 * i;
      long long c = (long long)q[0] | ((long long)q[1] << 16) | ((long long)q[2] << 32);
      c = c >= (1ll << 47) ? c - (1ll << 48) : c;
      m = (double)c;
    } else {
      m = (double)((const int*)Q)[i];
    }
    return m * inv[slot];
  }
  __device__ __forceinline__ double d1(int a) const { return get(3 * cur + a); }
  __device__ __forceinline__ double d2(int i, int j) const {
    return get(72 + 6 * s2slot(cur) + sym(i, j));
  }
  __device__ __forceinline__ double ko() const { return 0.0; }
};

template <bool WIDE>
__global__ __launch_bounds__(256) void k_algebra_bfp(const double* __restrict__ F,
                                                     const unsigned char* __restrict__ Q,
                                                     double* __restrict__ OUT, long n, Scal sc) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  BundleSrcQ<WIDE> s{F, (const double*)Q, Q + BFP_HEADER_BYTES, n, p, 0};
  PlainOut o{OUT, n, p};
  bssn_rhs_point<false>(s, with_ssl(sc, nullptr, 0.0), o);
}

hipError_t launch_algebra_bfp(const double* F, const void* Q, double* OUT, long npts,
                              const Scal& sc, bool wide, hipStream_t st) {
  const unsigned nb = (unsigned)((npts + 255) / 256);
  if (wide)
    hipLaunchKernelGGL(k_algebra_bfp<true>, dim3(nb), dim3(256), 0, st, F,
                       (const unsigned char*)Q, OUT, npts, sc);
  else
    hipLaunchKernelGGL(k_algebra_bfp<false>, dim3(nb), dim3(256), 0, st, F,
                       (const unsigned char*)Q, OUT, npts, sc);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_algebra(const double* __restrict__ F,
                                                 const double* __restrict__ D,
                                                 double* __restrict__ OUT, long n, Scal sc) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  BundleSrc s{F, D, n, p, 0};
  PlainOut o{OUT, n, p};
  bssn_rhs_point<false>(s, with_ssl(sc, nullptr, 0.0), o);
}

hipError_t launch_algebra(const double* F, const double* D, double* OUT, long npts,
                          const Scal& sc, hipStream_t st) {
  const unsigned nb = (unsigned)((npts + 255) / 256);
  hipLaunchKernelGGL(k_algebra, dim3(nb), dim3(256), 0, st, F, D, OUT, npts, sc);
  return hipGetLastError();
}

// ------------------------------------------------------------------------- KO pass
template <int ORD, int MODE>
__global__ __launch_bounds__(64) void k_ko_add(const double* __restrict__ F,
                                               double* __restrict__ OUT, Dims d, Scales w) {
  using G = Geo<ORD, true>;
  __shared__ double lds[G::VOL];
  int ox, oy, oz;
  if (!brick_origin<G>(d, &ox, &oy, &oz)) return;
  const int lane = (int)threadIdx.x;
  const bool valid = lane_valid<G>(d, ox, oy, oz, lane);
  // FD4: two fields in flight, stepped in pairs (constant staging slots); FD6 keeps one. The
  // next field's RHS value is loaded one iteration ahead (gc is clamped into the grid, so the
  // load is in bounds for every lane).
  constexpr int DEPTH = ORD == 4 ? 2 : 1;
  WaveSrc<ORD, true, MODE, StateOrder, DEPTH> s(lds, F, w, d, ox, oy, oz, lane);
  double o = OUT[s.gc];
  auto field = [&](int f, auto slot) {
    s.template stage_slot<decltype(slot)::value>(f);
    const double k = s.ko();
    const long i = (long)f * d.P + s.gc;
    if (valid) OUT[i] = o + k;
    if (f + 1 < BSSN_NF) o = OUT[i + d.P];
    __builtin_amdgcn_wave_barrier();
  };
  for_fields<DEPTH>(field);
}

// FD order 8 takes Kreiss-Oliger order 8, as FD order 6 does: the same kernel (the KO stencil
// and its scale sigma / (256 h) do not depend on the FD order beyond the pairing).
hipError_t launch_ko_add(const double* F, double* OUT, const Dims& d, const Scales& w, int ord,
                         int mode, hipStream_t st) {
  if (ord == 4) {
    if (mode == WRAP)
      hipLaunchKernelGGL((k_ko_add<4, WRAP>), dim3(brick_blocks<Geo<4, true>>(d)), dim3(64), 0,
                         st, F, OUT, d, w);
    else
      hipLaunchKernelGGL((k_ko_add<4, CLAMP>), dim3(brick_blocks<Geo<4, true>>(d)), dim3(64), 0,
                         st, F, OUT, d, w);
  } else {
    if (mode == WRAP)
      hipLaunchKernelGGL((k_ko_add<6, WRAP>), dim3(brick_blocks<Geo<6, true>>(d)), dim3(64), 0,
                         st, F, OUT, d, w);
    else
      hipLaunchKernelGGL((k_ko_add<6, CLAMP>), dim3(brick_blocks<Geo<6, true>>(d)), dim3(64), 0,
                         st, F, OUT, d, w);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------------- periodic sync
// Every ghost point takes the value of its periodic image in the interior. One pass suffices:
// the sources are interior points, which is also what the reference's x-then-y-then-z slab
// copies produce at edges and corners. One thread per ghost point (ghost_point).
__global__ __launch_bounds__(256) void k_periodic_sync(double* __restrict__ F, Dims d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ghost_count(d)) return;
  int c[3];
  ghost_point(d, t, c);
  const long p = ((long)c[0] * d.S[1] + c[1]) * d.S[2] + c[2];
  int s[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int N = d.S[a] - 2 * d.ng;
    int v = (c[a] - d.ng) % N;
    s[a] = d.ng + (v < 0 ? v + N : v);
  }
  const long q = ((long)s[0] * d.S[1] + s[1]) * d.S[2] + s[2];
  // all 24 loads before the stores (the source is interior, the destination a ghost), so they
  // are in flight together instead of each waiting behind the previous store
  double v[BSSN_NF];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) v[f] = F[(long)f * d.P + q];
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) F[(long)f * d.P + p] = v[f];
}

hipError_t launch_periodic_sync(double* F, const Dims& d, hipStream_t st) {
  if (d.ng <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((ghost_count(d) + 255) / 256);
  hipLaunchKernelGGL(k_periodic_sync, dim3(nb), dim3(256), 0, st, F, d);
  return hipGetLastError();
}

// ------------------------------------------------------------------------- enforce
__global__ __launch_bounds__(256) void k_enforce(double* __restrict__ F, long P, double chi_floor,
                                                 double alpha_floor) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double g[6], a[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    g[i] = F[(long)(GT0 + i) * P + p];
    a[i] = F[(long)(AT0 + i) * P + p];
  }
  const double det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
                     g[2] * (g[1] * g[4] - g[3] * g[2]);
  const double scale = pow(det, -1.0 / 3.0);
#pragma unroll
  for (int i = 0; i < 6; ++i) g[i] = g[i] * scale;
  const double iuxx = g[3] * g[5] - g[4] * g[4];
  const double iuxy = g[2] * g[4] - g[1] * g[5];
  const double iuxz = g[1] * g[4] - g[2] * g[3];
  const double iuyy = g[0] * g[5] - g[2] * g[2];
  const double iuyz = g[1] * g[2] - g[0] * g[4];
  const double iuzz = g[0] * g[3] - g[1] * g[1];
  const double trA = iuxx * a[0] + iuyy * a[3] + iuzz * a[5] +
                     2.0 * (iuxy * a[1] + iuxz * a[2] + iuyz * a[4]);
  const double third = trA / 3.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    F[(long)(GT0 + i) * P + p] = g[i];
    F[(long)(AT0 + i) * P + p] = a[i] - g[i] * third;
  }
  const double c = F[(long)CHI * P + p];
  F[(long)CHI * P + p] = c < chi_floor ? chi_floor : c;
  const double al = F[(long)ALPHA * P + p];
  F[(long)ALPHA * P + p] = al < alpha_floor ? alpha_floor : al;
}

hipError_t launch_enforce(double* F, long P, double chi_floor, double alpha_floor,
                          hipStream_t st) {
  const unsigned nb = (unsigned)((P + 255) / 256);
  hipLaunchKernelGGL(k_enforce, dim3(nb), dim3(256), 0, st, F, P, chi_floor, alpha_floor);
  return hipGetLastError();
}

// ------------------------------------------------------------------------- constraints
struct ConsOut {
  double* HM;
  long P, gp;
  bool valid;
  __device__ __forceinline__ void put(int c, double v) {
    if (valid) HM[(long)c * P + gp] = v;
  }
};

template <int ORD>
__global__ __launch_bounds__(64) void k_constraints(const double* __restrict__ F,
                                                    double* __restrict__ HM, Dims d, Scales w) {
  using G = Geo<ORD, false>;
  __shared__ double lds[G::VOL];
  int ox, oy, oz;
  if (!brick_origin<G>(d, &ox, &oy, &oz)) return;
  const int lane = (int)threadIdx.x;
  WaveSrc<ORD, false, CLAMP, ConsOrder> s(lds, F, w, d, ox, oy, oz, lane);
  ConsOut o{HM, d.P, s.gc, lane_valid<G>(d, ox, oy, oz, lane)};
  bssn_constraints_point(s, o);
}

hipError_t launch_constraints(const double* F, double* HM, const Dims& d, const Scales& w,
                              int ord, hipStream_t st) {
  with_order(ord, [&](auto oc) {
    constexpr int O = decltype(oc)::value;
    hipLaunchKernelGGL(k_constraints<O>, dim3(brick_blocks<Geo<O, false>>(d)), dim3(64), 0, st,
                       F, HM, d, w);
  });
  return hipGetLastError();
}

// Interior sums of H^2 and |M|^2: one (sum H^2, sum |M|^2) pair per block, in a fixed order,
// so the reduction is deterministic. Blocks cover (x, y) rows of the interior.
int sumsq_blocks(const Dims& d) { return d.n[0] * d.n[1]; }

__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ HM, Dims d,
                                               double* __restrict__ partial) {
  __shared__ double sh[256], sm[256];
  const int row = (int)blockIdx.x;
  const int x = d.lo[0] + row / d.n[1], y = d.lo[1] + row % d.n[1];
  double h = 0.0, m = 0.0;
  for (int k = (int)threadIdx.x; k < d.n[2]; k += 256) {
    const long p = ((long)x * d.S[1] + y) * d.S[2] + d.lo[2] + k;
    const double H = HM[p], M0 = HM[d.P + p], M1 = HM[2 * d.P + p], M2 = HM[3 * d.P + p];
    h += H * H;
    m += M0 * M0 + M1 * M1 + M2 * M2;
  }
  sh[threadIdx.x] = h;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x] += sh[threadIdx.x + s];
      sm[threadIdx.x] += sm[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * row] = sh[0];
    partial[2 * row + 1] = sm[0];
  }
}

hipError_t launch_sumsq(const double* HM, const Dims& d, double* partial, int nblk,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq, dim3((unsigned)nblk), dim3(256), 0, st, HM, d, partial);
  return hipGetLastError();
}

// total[0..1] = the partial sums added in row order from 0.0, by one thread: the same sums, in
// the same order, as the host loop of bssn_constraint_l2_f64, so both entry points agree bit
// for bit
__global__ void k_sum_partials(const double* __restrict__ partial, int nblk,
                               double* __restrict__ total) {
  if (threadIdx.x != 0) return;
  double sh = 0.0, sm = 0.0;
  for (int b = 0; b < nblk; ++b) {
    sh += partial[2 * b];
    sm += partial[2 * b + 1];
  }
  total[0] = sh;
  total[1] = sm;
}

hipError_t launch_sum_partials(const double* partial, int nblk, double* total, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial, nblk, total);
  return hipGetLastError();
}

// Static workspace of bssn_constraint_l2_static_f64 (for callers that cannot allocate): the
// four constraint arrays, the row partial sums and the two totals, BSSN_L2_SCRATCH_DOUBLES
// doubles (80 MiB) per device, held by every device the library runs on until the process
// ends. Host code uses the address hipGetSymbolAddress returns for the current device, never
// the array's name.
static __device__ __align__(256) double scratch[BSSN_L2_SCRATCH_DOUBLES];

hipError_t l2_scratch(double** out) {
  static std::mutex mu;
  static double* addr[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = (double*)p;
  }
  *out = addr[dev];
  return hipSuccess;
}

}  // namespace bssn
