// Shared host/device definitions for the BSSN HIP library (gfx950).
//
// State layout (the reference's, 3D/src/bssn3d/state.py:26-44 and gen_fused.py:210-211):
//   F[f*P + (x*Sy + y)*Sz + z], P = Sx*Sy*Sz, 24 fields in the order
//   alpha, chi, K, gt0..5, beta0..2, At0..5, Gt0..2, B0..2.
// Derivative bundle layout (two-pass path): D[d*P + p], d = 3*f + axis for the 72 first
// derivatives, d = 72 + 6*s + pair for the 66 second derivatives (s over alpha, chi, gt0..5,
// beta0..2; pair over 00,01,02,11,12,22).
#pragma once

#include <stddef.h>

#define BSSN_NF 24
#define BSSN_ND 138
#define BSSN_MAXTAP 9

// Brick of output points one wave owns (one point per lane): extents along x, y, z, product 64.
// A schedule knob of the autotuner; host brick lists and device kernels share it.
#ifndef BSSN_BRICK_X
#define BSSN_BRICK_X 4
#endif
#ifndef BSSN_BRICK_Y
#define BSSN_BRICK_Y 4
#endif
#ifndef BSSN_BRICK_Z
#define BSSN_BRICK_Z 4
#endif
static_assert(BSSN_BRICK_X * BSSN_BRICK_Y * BSSN_BRICK_Z == 64, "a brick is one 64-lane wave");

namespace bssn {

// Halo handling of the stencil source (see bssn_brick.h).
// LINK (batched blocks only): cells that are ghosts of the block are read where the link map
// points (a neighbouring block's interior point, or the block's own ghost where a prolongation
// fills it), so a step needs no ghost copies between its stages.
enum Mode : int { CLAMP = 0, WRAP = 1, LINK = 2 };

enum Field : int {
  ALPHA = 0, CHI = 1, KTR = 2,
  GT0 = 3, GT1 = 4, GT2 = 5, GT3 = 6, GT4 = 7, GT5 = 8,
  BETA0 = 9, BETA1 = 10, BETA2 = 11,
  AT0 = 12, AT1 = 13, AT2 = 14, AT3 = 15, AT4 = 16, AT5 = 17,
  GTI0 = 18, GTI1 = 19, GTI2 = 20,
  BB0 = 21, BB1 = 22, BB2 = 23
};

// lane -> coordinates within the brick (z fastest), for the kernels that walk brick lists
constexpr int brick_lx(int lane) { return lane / (BSSN_BRICK_Y * BSSN_BRICK_Z); }
constexpr int brick_ly(int lane) { return (lane / BSSN_BRICK_Z) % BSSN_BRICK_Y; }
constexpr int brick_lz(int lane) { return lane % BSSN_BRICK_Z; }

// Physics scalars, unpacked from the reference's 16-double pack
// (eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz;
// 3D/src/bssn3d/fused_rhs_cuda.py:52-55), plus two folded factors.
struct Scal {
  double eta;
  double lam[4];
  double lamf[2];
  double ssl_h;        // h_ssl
  double ssl_t;        // evaluation time t
  double ssl_sig;      // sig_ssl
  double ssl_e;        // exp(-t^2 / (2 sig_ssl^2)): always formed on the device (with_ssl),
                       // so host-pack, device-pack and device-time launches agree bit for bit
  double cahd;         // C_CAHD * dx_i^2 / (dt * (10 dx_i^2 + 1))
};

// Offsets into the 16-double scalar pack.
enum PackSlot : int {
  PK_ETA = 0, PK_CAHD = 1, PK_DT = 2, PK_DXI = 3, PK_SSL_H = 4, PK_SSL_SIG = 5, PK_T = 6,
  PK_LAM = 7, PK_LAMF = 11, PK_H = 13
};

// Per-axis scales of the integer stencils: the centred coefficients are applied as their
// integer numerators and the result is multiplied once by these.
struct Scales {
  double s1[3];     // 1 / (den1 h_a)
  double s2[3];     // 1 / (den2 h_a^2)
  double sm[3];     // mixed (01, 02, 12): 1 / (den1^2 h_i h_j)
  double sko[3];    // sigma / (den_ko h_a)
};

struct Dims {
  int S[3];      // padded extents Sx, Sy, Sz
  int lo[3];     // first output point per axis
  int n[3];      // output points per axis
  int ng;        // ghost width (wrap mode: interior = [ng, S - ng))
  long P;        // Sx*Sy*Sz
};

#if defined(__HIPCC__)
#include <math.h>

// Scales of the integer stencils: the denominators of the centred and Kreiss-Oliger
// coefficients (common/src/mcs_common/derivatives.py:33-49) times the spacings. Computed on
// the host for one grid and on the device per block of a batch; contraction is off so that
// both round identically (division and multiplication are correctly rounded on both).
__host__ __device__ inline Scales make_scales(const double h3[3], int ord, double sigma) {
#pragma clang fp contract(off)
  Scales s;
  const double den1 = ord == 4 ? 12.0 : (ord == 6 ? 60.0 : 840.0);     // first derivative
  const double den2 = ord == 4 ? 12.0 : (ord == 6 ? 180.0 : 5040.0);    // second derivative
  const double denk = ord == 4 ? 64.0 : 256.0;   // KO of order min(ord + 2, 8)
  for (int a = 0; a < 3; ++a) {
    s.s1[a] = 1.0 / (den1 * h3[a]);
    s.s2[a] = 1.0 / (den2 * (h3[a] * h3[a]));
    s.sko[a] = sigma / (denk * h3[a]);
  }
  s.sm[0] = 1.0 / (den1 * den1 * (h3[0] * h3[1]));
  s.sm[1] = 1.0 / (den1 * den1 * (h3[0] * h3[2]));
  s.sm[2] = 1.0 / (den1 * den1 * (h3[1] * h3[2]));
  return s;
}

// CAHD factor C dx_i^2 / (dt (10 dx_i^2 + 1)) (vendor/bssneqs_SSL_HD_dxsq.cpp:857), same
// rounding on host and device.
__host__ __device__ inline double cahd_factor(double C, double dx, double dt) {
#pragma clang fp contract(off)
  const double dx2 = dx * dx;
  return C * dx2 / (dt * (10.0 * dx2 + 1.0));
}

// A wave-uniform double computed by vector instructions (there is no scalar FP64 arithmetic),
// moved to scalar registers so that it does not occupy a vector register for the whole kernel.
__device__ __forceinline__ double uniform(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The SSL factor exp(-t^2 / (2 sig^2)), formed on the device at the evaluation time: sc.ssl_t,
// or a stage time held in device memory (*tdev + toff: a captured step replayed with its
// time advanced on the device).
__device__ __forceinline__ Scal with_ssl(Scal sc, const double* tdev, double toff) {
  const double t = tdev ? *tdev + toff : sc.ssl_t;
  sc.ssl_e = uniform(exp(-0.5 * t * t / (sc.ssl_sig * sc.ssl_sig)));
  return sc;
}

// Scalars of one evaluation from a pack S held in device memory (the reference's Fused(stream,
// F, S) passes S as a device buffer): the same values, in the same rounding, as the host's
// unpacking of S at time t (contraction off).
__device__ __forceinline__ Scal scal_from_pack(const double* S, double dt, double t) {
#pragma clang fp contract(off)
  Scal sc;
  sc.eta = uniform(S[PK_ETA]);
#pragma unroll
  for (int i = 0; i < 4; ++i) sc.lam[i] = uniform(S[PK_LAM + i]);
  sc.lamf[0] = uniform(S[PK_LAMF]);
  sc.lamf[1] = uniform(S[PK_LAMF + 1]);
  sc.ssl_h = uniform(S[PK_SSL_H]);
  sc.ssl_sig = uniform(S[PK_SSL_SIG]);
  sc.ssl_t = uniform(t);
  sc.ssl_e = 0.0;
  sc.cahd = uniform(cahd_factor(S[PK_CAHD], S[PK_DXI], dt));
  return sc;
}

// Stencil scales from a device pack (spacings S[13..15]), moved to scalar registers.
__device__ __forceinline__ Scales scales_from_pack(const double* S, int ord, double sigma) {
  const double h3[3] = {S[PK_H], S[PK_H + 1], S[PK_H + 2]};
  Scales w = make_scales(h3, ord, sigma);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    w.s1[a] = uniform(w.s1[a]);
    w.s2[a] = uniform(w.s2[a]);
    w.sm[a] = uniform(w.sm[a]);
    w.sko[a] = uniform(w.sko[a]);
  }
  return w;
}

// Keep a value where it is computed: an empty volatile asm that claims to rewrite x. The
// compiler must evaluate x before this point and may not sink its arithmetic past later
// barriers, which otherwise keeps raw stencil taps live across field stages.
__device__ __forceinline__ double pin(double x) {
  asm volatile("" : "+v"(x));
  return x;
}
#endif

}  // namespace bssn
