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

namespace bssn {

// Halo handling of the stencil source (see bssn_brick.h).
enum Mode : int { CLAMP = 0, WRAP = 1 };

enum Field : int {
  ALPHA = 0, CHI = 1, KTR = 2,
  GT0 = 3, GT1 = 4, GT2 = 5, GT3 = 6, GT4 = 7, GT5 = 8,
  BETA0 = 9, BETA1 = 10, BETA2 = 11,
  AT0 = 12, AT1 = 13, AT2 = 14, AT3 = 15, AT4 = 16, AT5 = 17,
  GTI0 = 18, GTI1 = 19, GTI2 = 20,
  BB0 = 21, BB1 = 22, BB2 = 23
};

// Physics scalars, unpacked from the reference's 16-double pack
// (eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz;
// 3D/src/bssn3d/fused_rhs_cuda.py:52-55), plus two host-folded factors.
struct Scal {
  double eta;
  double lam[4];
  double lamf[2];
  double ssl_h;        // h_ssl
  double ssl_e;        // exp(-t^2 / (2 sig_ssl^2)) at the evaluation time
  double cahd;         // C_CAHD * dx_i^2 / (dt * (10 dx_i^2 + 1))
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
// Keep a value where it is computed: an empty volatile asm that claims to rewrite x. The
// compiler must evaluate x before this point and may not sink its arithmetic past later
// barriers, which otherwise keeps raw stencil taps live across field stages.
__device__ __forceinline__ double pin(double x) {
  asm volatile("" : "+v"(x));
  return x;
}
#endif

}  // namespace bssn
