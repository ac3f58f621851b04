// Wave-private LDS bricks for the fused kernels.
//
// One workgroup is one wave of 64 lanes owning a 4x4x4 brick of output points (one point per
// lane, z fastest). The 24 fields are staged one at a time into a private LDS box of edge
// 4 + 2H; since a wave's LDS operations execute in issue order, staging needs no workgroup
// barrier, so every wave runs independently and latency is hidden across the waves of a CU.
//
// Only the cells a stencil can reach are staged: the core, the six face slabs to depth H
// (first/second derivatives and Kreiss-Oliger along one axis) and the twelve edge bars to depth
// R in two axes (mixed derivatives). For FD order 4 with KO order 6 that is 544 of the 1000
// cells of the box, 8.5 loads per point and field.
//
// Halo positions are mapped to the grid once per lane (not per field):
//   CLAMP  x -> min(max(x, 0), S-1) per axis. The reference edge-pads each one-axis operator by
//          its own reach and forms mixed derivatives as d1_j(d1_i f)
//          (common/src/mcs_common/derivatives.py:78-88, derivative_bundle.py:41-46); clamping is
//          separable, so stencils over the clamped box equal those exactly, ghosts included.
//   WRAP   x -> ng + ((x - ng) mod N) per axis (N = S - 2 ng): the values a periodic ghost sync
//          (3D/src/bssn3d/evolve.py:65-73) would have put there, so no sync pass is needed
//          between RK stages.
//
// Stencils use the integer numerators of the centred coefficients
// (common/src/mcs_common/derivatives.py:33-49), paired symmetrically, with one scale per axis
// (1/(den h), 1/(den h^2), 1/(den^2 h_i h_j), sigma/(den h)) applied once.
#pragma once

#include <hip/hip_runtime.h>

#include "bssn_params.h"

namespace bssn {

// Integer stencil numerators by reach k = 0..R (d1: antisymmetric, index 0 unused).
template <int ORD>
struct Coef;
template <>
struct Coef<4> {
  static constexpr int R = 2;
  __device__ static constexpr double d1(int k) { return k == 1 ? 8.0 : (k == 2 ? -1.0 : 0.0); }
  __device__ static constexpr double d2(int k) {
    return k == 0 ? -30.0 : (k == 1 ? 16.0 : -1.0);
  }
  // Kreiss-Oliger of order 6: [1, -6, 15, -20, 15, -6, 1] / 64
  static constexpr int RK = 3;
  __device__ static constexpr double ko(int k) {
    return k == 0 ? -20.0 : (k == 1 ? 15.0 : (k == 2 ? -6.0 : 1.0));
  }
};
template <>
struct Coef<6> {
  static constexpr int R = 3;
  __device__ static constexpr double d1(int k) {
    return k == 1 ? 45.0 : (k == 2 ? -9.0 : (k == 3 ? 1.0 : 0.0));
  }
  __device__ static constexpr double d2(int k) {
    return k == 0 ? -490.0 : (k == 1 ? 270.0 : (k == 2 ? -27.0 : 2.0));
  }
  // Kreiss-Oliger of order 8: [-1, 8, -28, 56, -70, 56, -28, 8, -1] / 256
  static constexpr int RK = 4;
  __device__ static constexpr double ko(int k) {
    return k == 0 ? -70.0 : (k == 1 ? 56.0 : (k == 2 ? -28.0 : (k == 3 ? 8.0 : -1.0)));
  }
};

template <int ORD, bool KO>
struct Geo {
  static constexpr int R = Coef<ORD>::R;                  // FD reach
  static constexpr int RK = KO ? Coef<ORD>::RK : 0;       // KO reach
  static constexpr int H = R > RK ? R : RK;               // box halo
  static constexpr int B = 4;                             // brick edge
  static constexpr int NT = 64;                           // lanes (one wave)
  static constexpr int N = B + 2 * H;                     // box edge
  static constexpr int SX = N * N, SY = N;                // LDS strides (z contiguous)
  static constexpr int VOL = N * N * N;
  static constexpr int NCORE = B * B * B;
  static constexpr int NFACE = 6 * B * B * H;
  static constexpr int NEDGE = 12 * B * R * R;
  static constexpr int NCELL = NCORE + NFACE + NEDGE;
  static constexpr int PER = (NCELL + NT - 1) / NT;       // staged cells per lane
};

// Box coordinates of staged cell s (core, then faces, then edges); s >= NCELL repeats cell 0.
template <class G>
__device__ __forceinline__ void cell_of_slot(int s, int* c) {
  constexpr int B = G::B, H = G::H, R = G::R;
  if (s >= G::NCELL) s = 0;
  if (s < G::NCORE) {
    c[0] = H + s / (B * B);
    c[1] = H + (s / B) % B;
    c[2] = H + s % B;
    return;
  }
  s -= G::NCORE;
  if (s < G::NFACE) {
    const int a = s / (2 * B * B * H);
    int r = s % (2 * B * B * H);
    const int side = r / (B * B * H);
    r %= B * B * H;
    const int depth = r / (B * B);
    const int q = r % (B * B);
    const int u = H + q / B, v = H + q % B;
    const int w = side == 0 ? H - 1 - depth : H + B + depth;
    // a = axis normal to the face; (u, v) run over the other two axes in order
    if (a == 0) { c[0] = w; c[1] = u; c[2] = v; }
    else if (a == 1) { c[0] = u; c[1] = w; c[2] = v; }
    else { c[0] = u; c[1] = v; c[2] = w; }
    return;
  }
  s -= G::NFACE;
  // 12 edge bars: axis pair p (01, 02, 12) x 4 sign combinations; the third axis is core
  const int e = s / (B * R * R);
  int r = s % (B * R * R);
  const int p = e / 4, sg = e % 4;
  const int t = H + r / (R * R);
  r %= R * R;
  const int di = r / R, dj = r % R;
  const int wi = (sg & 1) ? H + B + di : H - 1 - di;
  const int wj = (sg & 2) ? H + B + dj : H - 1 - dj;
  if (p == 0) { c[0] = wi; c[1] = wj; c[2] = t; }
  else if (p == 1) { c[0] = wi; c[1] = t; c[2] = wj; }
  else { c[0] = t; c[1] = wi; c[2] = wj; }
}

// Source of fields and derivatives for bssn_rhs_point: one wave-private LDS box per field.
template <int ORD, bool KO, int MODE>
struct WaveSrc {
  using G = Geo<ORD, KO>;
  using C = Coef<ORD>;
  double* lds;          // VOL doubles, this wave's box
  const double* F;      // stencil source state
  const Scales& sc;
  long P;
  int off[G::PER];      // grid offsets of this lane's staged cells
  int cel[G::PER];      // box indices of this lane's staged cells
  long gc;              // flat index of the lane's point (clamped into the grid)
  int c;                // box index of the lane's point

  __device__ __forceinline__ WaveSrc(double* lds_, const double* F_, const Scales& sc_,
                                     const Dims& d, int ox, int oy, int oz, int lane)
      : lds(lds_), F(F_), sc(sc_), P(d.P) {
    const int tx = lane / (G::B * G::B), ty = (lane / G::B) % G::B, tz = lane % G::B;
    c = ((tx + G::H) * G::N + (ty + G::H)) * G::N + (tz + G::H);
    int p[3] = {ox + tx, oy + ty, oz + tz};
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = p[a] > d.S[a] - 1 ? d.S[a] - 1 : p[a];
    gc = ((long)p[0] * d.S[1] + p[1]) * d.S[2] + p[2];
    const int o[3] = {ox, oy, oz};
#pragma unroll
    for (int k = 0; k < G::PER; ++k) {
      int cc[3];
      cell_of_slot<G>(lane + k * G::NT, cc);
      cel[k] = (cc[0] * G::N + cc[1]) * G::N + cc[2];
      int g[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int v = o[a] + cc[a] - G::H;
        if (MODE == CLAMP) {
          v = v < 0 ? 0 : (v > d.S[a] - 1 ? d.S[a] - 1 : v);
        } else {
          const int N = d.S[a] - 2 * d.ng;
          int w = (v - d.ng) % N;   // once per lane and slot, not per field
          w = w < 0 ? w + N : w;
          v = d.ng + w;
        }
        g[a] = v;
      }
      off[k] = (g[0] * d.S[1] + g[1]) * d.S[2] + g[2];
    }
  }

  __device__ __forceinline__ void stage(int f) {
    const double* u = F + (long)f * P;
    double v[G::PER];
#pragma unroll
    for (int k = 0; k < G::PER; ++k) v[k] = u[off[k]];
#pragma unroll
    for (int k = 0; k < G::PER; ++k) lds[cel[k]] = v[k];
    // The wave's LDS accesses complete in issue order; this only stops the compiler from
    // moving this field's reads above the writes.
    __builtin_amdgcn_wave_barrier();
  }

  __device__ __forceinline__ double center(int f) const { return F[(long)f * P + gc]; }
  __device__ __forceinline__ double val() const { return pin(lds[c]); }

  __device__ __forceinline__ static constexpr int stride(int a) {
    return a == 0 ? G::SX : (a == 1 ? G::SY : 1);
  }

  // unscaled first-derivative numerator along axis a at box index q
  __device__ __forceinline__ double n1(int q, int a) const {
    const int s = stride(a);
    double acc = 0.0;
#pragma unroll
    for (int k = 1; k <= G::R; ++k) acc += C::d1(k) * (lds[q + k * s] - lds[q - k * s]);
    return acc;
  }

  __device__ __forceinline__ double d1(int a) const { return pin(n1(c, a) * sc.s1[a]); }

  __device__ __forceinline__ double d2(int i, int j) const {
    if (i == j) {
      const int s = stride(i);
      double acc = C::d2(0) * lds[c];
#pragma unroll
      for (int k = 1; k <= G::R; ++k) acc += C::d2(k) * (lds[c + k * s] + lds[c - k * s]);
      return pin(acc * sc.s2[i]);
    }
    const int sj = stride(j);
    double acc = 0.0;
#pragma unroll
    for (int k = 1; k <= G::R; ++k) acc += C::d1(k) * (n1(c + k * sj, i) - n1(c - k * sj, i));
    return pin(acc * sc.sm[i + j - 1]);
  }

  __device__ __forceinline__ double ko() const {
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int s = stride(a);
      double t = C::ko(0) * lds[c];
#pragma unroll
      for (int k = 1; k <= G::RK; ++k) t += C::ko(k) * (lds[c + k * s] + lds[c - k * s]);
      acc += t * sc.sko[a];
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

template <class G>
__device__ __forceinline__ bool brick_origin(const Dims& d, int* ox, int* oy, int* oz) {
  const int nbx = (d.n[0] + G::B - 1) / G::B;
  const int nby = (d.n[1] + G::B - 1) / G::B;
  const int nbz = (d.n[2] + G::B - 1) / G::B;
  int bx, by, bz;
  if (!brick_of_block(nbx, nby, nbz, &bx, &by, &bz)) return false;
  *ox = d.lo[0] + bx * G::B;
  *oy = d.lo[1] + by * G::B;
  *oz = d.lo[2] + bz * G::B;
  return true;
}

template <class G>
inline unsigned brick_blocks(const Dims& d) {
  const long nb = (long)((d.n[0] + G::B - 1) / G::B) * ((d.n[1] + G::B - 1) / G::B) *
                  ((d.n[2] + G::B - 1) / G::B);
  return (unsigned)((nb + 7) & ~7L);
}

template <class G>
__device__ __forceinline__ bool lane_valid(const Dims& d, int ox, int oy, int oz, int lane) {
  const int tx = lane / (G::B * G::B), ty = (lane / G::B) % G::B, tz = lane % G::B;
  return (ox + tx < d.lo[0] + d.n[0]) && (oy + ty < d.lo[1] + d.n[1]) &&
         (oz + tz < d.lo[2] + d.n[2]);
}

}  // namespace bssn
