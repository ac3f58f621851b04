// Wave-private LDS bricks for the fused kernels.
//
// One workgroup is one wave of 64 lanes owning a brick of output points (one point per lane,
// z fastest): 4x4x4 by default, any BSSN_BRICK_X x BSSN_BRICK_Y x BSSN_BRICK_Z of 64 points.
// The 24 fields are staged one at a time into a private LDS box (the brick plus H on every
// side); since a wave's LDS operations execute in issue order, staging needs no workgroup
// barrier, so every wave runs independently and latency is hidden across the waves of a CU.
//
// Only the cells a stencil can reach are staged: the core, the six face slabs to depth H
// (first/second derivatives and Kreiss-Oliger along one axis) and the twelve edge bars to depth
// R in two axes (mixed derivatives). For the 4x4x4 brick, FD order 4 with KO order 6, that is
// 544 of the 1000 cells of the box, 8.5 loads per point and field; the 13 fields without mixed derivatives
// (Gt^i, At_ij, K, B^i) skip the edge bars: 352 cells, 5.5 loads per point.
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

#include "bssn_entries.h"
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
template <>
struct Coef<8> {
  static constexpr int R = 4;
  __device__ static constexpr double d1(int k) {
    return k == 1 ? 672.0 : (k == 2 ? -168.0 : (k == 3 ? 32.0 : (k == 4 ? -3.0 : 0.0)));
  }
  __device__ static constexpr double d2(int k) {
    return k == 0 ? -14350.0
         : (k == 1 ? 8064.0 : (k == 2 ? -1008.0 : (k == 3 ? 128.0 : -9.0)));
  }
  // Kreiss-Oliger of order 8 as well: the reference's KO stencils stop at order 8, and FD order
  // 8 takes it by default (common/src/mcs_common/derivatives.py:45-49, 64)
  static constexpr int RK = 4;
  __device__ static constexpr double ko(int k) { return Coef<6>::ko(k); }
};

// LDS box x-plane stride modulo 32 doubles (-1: dense planes); a schedule knob of the autotuner
#ifndef BSSN_BOX_XPLANE_MOD32
#define BSSN_BOX_XPLANE_MOD32 -1
#endif

template <int ORD, bool KO>
struct Geo {
  static constexpr int R = Coef<ORD>::R;                  // FD reach
  static constexpr int RK = KO ? Coef<ORD>::RK : 0;       // KO reach
  static constexpr int H = R > RK ? R : RK;               // box halo
  static constexpr int BX = BSSN_BRICK_X, BY = BSSN_BRICK_Y, BZ = BSSN_BRICK_Z;  // brick
  static constexpr int NT = 64;                           // lanes (one wave)
  static constexpr int NX = BX + 2 * H, NY = BY + 2 * H, NZ = BZ + 2 * H;   // box extents
  // z rows are padded to a pitch of 12 doubles (the 4x4x4 brick): for the 16 lanes of one LDS
  // access phase (4 y-rows x 4 z of the brick) the addresses ty*12 + tz fall on 16 distinct
  // bank pairs. Other bricks use the smallest even pitch that holds a row.
  static constexpr int PZ = NZ <= 12 ? 12 : NZ + (NZ & 1);
  // x planes: dense (NY * PZ), or, with BSSN_BOX_XPLANE_MOD32 = r, padded to the next stride
  // congruent to r modulo 32 doubles: r = 16 puts the two x-planes of the 4x4x4 brick that 32
  // lanes span (tx, tx + 1) on disjoint halves of 64 banks.
  static constexpr int SX0 = NY * PZ;
  static constexpr int SX = BSSN_BOX_XPLANE_MOD32 < 0
                                ? SX0
                                : SX0 + ((BSSN_BOX_XPLANE_MOD32 - SX0 % 32) % 32 + 32) % 32;
  static constexpr int SY = PZ;                           // LDS strides (z contiguous)
  static constexpr int VOL = NX * SX;
  static_assert(NX < 32 && NY < 32 && NZ < 32, "box coordinates are packed in 5 bits");
  static constexpr int NCORE = BX * BY * BZ;
  static constexpr int NFACE = 2 * H * (BY * BZ + BX * BZ + BX * BY);
  static constexpr int NEDGE = 4 * R * R * (BX + BY + BZ);
  // lane -> brick coordinates (z fastest)
  __host__ __device__ static constexpr int lx(int lane) { return brick_lx(lane); }
  __host__ __device__ static constexpr int ly(int lane) { return brick_ly(lane); }
  __host__ __device__ static constexpr int lz(int lane) { return brick_lz(lane); }
  static constexpr int NCELL = NCORE + NFACE + NEDGE;
  static constexpr int PER = (NCELL + NT - 1) / NT;       // staged cells per lane
  // fields without mixed derivatives need the core and the face slabs only: the first
  // NCORE + NFACE cells of the table
  static constexpr int PER_FACES = (NCORE + NFACE + NT - 1) / NT;
};

// The fields whose second derivatives the RHS, the bundle and the constraints form (so whose
// mixed derivatives read the edge bars of the box): alpha, chi, gt_ij, beta^i. The others are
// staged over the core and the face slabs only (Kreiss-Oliger and the one-axis stencils).
__host__ __device__ constexpr bool needs_edges(int f) {
  return f == ALPHA || f == CHI || (f >= GT0 && f <= GT5) || (f >= BETA0 && f <= BETA2);
}

// The staged cells, ordered row by row along z so that the lanes of one load instruction
// cover contiguous runs of a grid row: the 16 core rows (z over the whole box) and the x- and
// y-face rows (z over the core) -- the NCORE + NFACE cells every field needs -- then the edge
// bars of the mixed derivatives: the z extensions of the face rows within reach R of the core
// and the xy edge rows (z over the core). Entry = packed box coordinates x | y << 5 | z << 10;
// slots past the cell count repeat the first core cell.
template <class G>
struct CellTable {
  short c[G::PER * G::NT];
  constexpr CellTable() : c() {
    constexpr int BX = G::BX, BY = G::BY, BZ = G::BZ, H = G::H, R = G::R, NZ = G::NZ;
    int n = 0;
    auto row = [&](int x, int y, int z0, int z1) {
      for (int z = z0; z < z1; ++z) c[n++] = (short)(x | (y << 5) | (z << 10));
    };
    for (int x = H; x < H + BX; ++x)
      for (int y = H; y < H + BY; ++y) row(x, y, 0, NZ);
    // face rows over the core, then (second pass) their z extensions within reach R
    for (int pass = 0; pass < 2; ++pass)
      for (int ax = 0; ax < 2; ++ax)
        for (int side = 0; side < 2; ++side)
          for (int d = 1; d <= H; ++d)
            for (int u = H; u < H + (ax == 0 ? BY : BX); ++u) {
              const int w = side == 0 ? H - d : H + (ax == 0 ? BX : BY) - 1 + d;
              const int x = ax == 0 ? w : u, y = ax == 0 ? u : w;
              if (pass == 0) {
                row(x, y, H, H + BZ);
              } else if (d <= R) {
                row(x, y, H - R, H);
                row(x, y, H + BZ, H + BZ + R);
              }
            }
    for (int sx = 0; sx < 2; ++sx)
      for (int sy = 0; sy < 2; ++sy)
        for (int dx = 1; dx <= R; ++dx)
          for (int dy = 1; dy <= R; ++dy)
            row(sx == 0 ? H - dx : H + BX - 1 + dx, sy == 0 ? H - dy : H + BY - 1 + dy, H,
                H + BZ);
    while (n < G::PER * G::NT) c[n++] = (short)(H | (H << 5) | (H << 10));
  }
};

template <class G>
__constant__ CellTable<G> kCells = CellTable<G>();

// Field staging orders: the fused RHS stages in the order of bssn_rhs_point's phases; the
// per-field loops (bundle, KO) in state order.
struct RhsOrder {
  // gt0..5 | chi | Gt0..2 | alpha | beta0..2 | At0..5 | K | B0..2
  static constexpr int N = BSSN_NF;
  __host__ __device__ static constexpr int at(int i) {
    return i < 6 ? GT0 + i
         : i == 6 ? CHI
         : i < 10 ? GTI0 + (i - 7)
         : i == 10 ? ALPHA
         : i < 14 ? BETA0 + (i - 11)
         : i < 20 ? AT0 + (i - 14)
         : i == 20 ? KTR
         : BB0 + (i - 21);
  }
  __host__ __device__ static constexpr int pos(int f) {
    return f >= GT0 && f <= GT5 ? f - GT0
         : f == CHI ? 6
         : f >= GTI0 && f <= GTI0 + 2 ? 7 + f - GTI0
         : f == ALPHA ? 10
         : f >= BETA0 && f <= BETA2 ? 11 + f - BETA0
         : f >= AT0 && f <= AT5 ? 14 + f - AT0
         : f == KTR ? 20
         : 21 + f - BB0;
  }
};
struct StateOrder {
  static constexpr int N = BSSN_NF;
  __host__ __device__ static constexpr int at(int i) { return i; }
  __host__ __device__ static constexpr int pos(int f) { return f; }
};
// the constraints: gt0..5 | chi | Gt0..2 | At0..5 | K (pos = N for the fields they do not use)
struct ConsOrder {
  static constexpr int N = 17;
  __host__ __device__ static constexpr int at(int i) {
    return i < 6 ? GT0 + i : (i == 6 ? CHI : (i < 10 ? GTI0 + (i - 7) : (i < 16 ? AT0 + (i - 10)
                                                                                  : KTR)));
  }
  __host__ __device__ static constexpr int pos(int f) {
    return f >= GT0 && f <= GT5 ? f - GT0
         : f == CHI ? 6
         : f >= GTI0 && f <= GTI0 + 2 ? 7 + f - GTI0
         : f >= AT0 && f <= AT5 ? 10 + f - AT0
         : f == KTR ? 16
         : N;
  }
};

// flat index of the lane's point, clamped into the grid (write-back index of invalid lanes)
template <class G>
__device__ __forceinline__ long lane_point(const Dims& d, int ox, int oy, int oz, int lane) {
  const int tx = G::lx(lane), ty = G::ly(lane), tz = G::lz(lane);
  int p[3] = {ox + tx, oy + ty, oz + tz};
#pragma unroll
  for (int a = 0; a < 3; ++a) p[a] = p[a] > d.S[a] - 1 ? d.S[a] - 1 : p[a];
  return ((long)p[0] * d.S[1] + p[1]) * d.S[2] + p[2];
}

// Stencils on a staged box: the derivative interface of bssn_rhs_point over the LDS box `lds`
// (row-major, strides SX, SY, 1) at the lane's box index c.
template <class G, class C>
struct BoxStencil {
  double* lds;          // the current field's box
  const Scales& sc;
  int c;                // box index of the lane's point

  __device__ __forceinline__ BoxStencil(double* lds_, const Scales& sc_) : lds(lds_), sc(sc_) {}

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
    for (int k = 1; k <= G::R; ++k) {
      acc += C::d1(k) * (n1(c + k * sj, i) - n1(c - k * sj, i));
      // order 8: 64 box reads per mixed derivative; issued all at once they hold 128 registers
      // and the stage kernels spill (up to 404 B per lane). One tap pair at a time: RK4 step
      // at 128^3 8.49 -> 7.39 ms. The sums and their order are unchanged. Orders 4 and 6 do
      // not spill and run slower with the barrier (256^3 steps +1.2 % and +0.9 %).
      if constexpr (G::R >= 4) __builtin_amdgcn_sched_barrier(0);
    }
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

// halo position -> grid index along one axis (clamped, or wrapped over the interior)
template <int MODE>
__device__ __forceinline__ int map_axis(int v, int S, int ng) {
  if (MODE == CLAMP) return v < 0 ? 0 : (v > S - 1 ? S - 1 : v);
  const int N = S - 2 * ng;
  int w = (v - ng) % N;
  w = w < 0 ? w + N : w;
  return ng + w;
}

// Source of fields and derivatives for bssn_rhs_point: one wave-private LDS box per field.
// Staging is pipelined one field ahead: stage(f) writes field f (loaded during the previous
// stage) into the box, then issues the loads of the next field in ORDER, which complete while
// field f's derivatives and algebra run (DEPTH fields ahead: DEPTH > 1 needs the stage calls
// unrolled, so that each slot index is a constant). Contract: a kernel stages its fields exactly in the
// sequence of its ORDER (RhsOrder for bssn_rhs_point, ConsOrder for the constraints,
// StateOrder for per-field loops); a different sequence would stage the wrong field.
template <int ORD, bool KO, int MODE, class ORDER = RhsOrder, int DEPTH = 1>
struct WaveSrc : BoxStencil<Geo<ORD, KO>, Coef<ORD>> {
  using G = Geo<ORD, KO>;
  using C = Coef<ORD>;
  using BoxStencil<G, C>::lds;
  using BoxStencil<G, C>::c;
  const double* F;      // stencil source state (LINK: the pool)
  const double* Fc;     // the block's state (centre values)
  long P;
  int off[G::PER];      // grid offsets of this lane's staged cells (LINK: pool offsets)
  int psrc[MODE == LINK ? G::PER : 1];   // LINK: field stride of each staged cell's block
  int cel[G::PER];      // box indices of this lane's staged cells
  double pf[DEPTH][G::PER];   // the next DEPTH fields' cells, in flight (slot = position % DEPTH)
  long gc;              // flat index of the lane's point (clamped into the grid)
  static constexpr bool kEarlyCentres = true;   // bssn_rhs_point: prefetch centres

  __device__ __forceinline__ WaveSrc(double* lds_, const double* F_, const Scales& sc_,
                                     const Dims& d, int ox, int oy, int oz, int lane)
      : BoxStencil<G, C>(lds_, sc_), F(F_), Fc(F_), P(d.P) {
    static_assert(MODE != LINK, "LINK sources take the pool and the link map");
    const int tx = G::lx(lane), ty = G::ly(lane), tz = G::lz(lane);
    c = (tx + G::H) * G::SX + (ty + G::H) * G::SY + (tz + G::H);
    gc = lane_point<G>(d, ox, oy, oz, lane);
    const int o[3] = {ox, oy, oz};
#pragma unroll
    for (int k = 0; k < G::PER; ++k) {
      const int e = kCells<G>.c[lane + k * G::NT];
      const int cc[3] = {e & 31, (e >> 5) & 31, (e >> 10) & 31};
      cel[k] = cc[0] * G::SX + cc[1] * G::SY + cc[2];
      int g[3];   // mapped once per lane and slot, not per field
#pragma unroll
      for (int a = 0; a < 3; ++a) g[a] = map_axis<MODE>(o[a] + cc[a] - G::H, d.S[a], d.ng);
      off[k] = (g[0] * d.S[1] + g[1]) * d.S[2] + g[2];
    }
#pragma unroll
    for (int i = 0; i < DEPTH; ++i)
      if (i < ORDER::N) fetch(i, ORDER::at(i));
  }

  // LINK: a block of a batch (pool offset boff, link map lmap of its padded points). Cells in
  // the block's interior are read in place; a ghost cell is read where the map points, or in
  // place if the entry would reach outside the pool (pool_n doubles). Needs ng >= H (the box
  // stays inside the padded block) and a pool below 2^29 doubles.
  __device__ __forceinline__ WaveSrc(double* lds_, const double* pool, long boff,
                                     const int* lmap, long long pool_n, const Scales& sc_,
                                     const Dims& d, int ox, int oy, int oz, int lane)
      : BoxStencil<G, C>(lds_, sc_), F(pool), Fc(pool + boff), P(d.P) {
    static_assert(MODE == LINK, "the link constructor is for LINK sources");
    const int tx = G::lx(lane), ty = G::ly(lane), tz = G::lz(lane);
    c = (tx + G::H) * G::SX + (ty + G::H) * G::SY + (tz + G::H);
    gc = lane_point<G>(d, ox, oy, oz, lane);
    const int o[3] = {ox, oy, oz};
#pragma unroll
    for (int k = 0; k < G::PER; ++k) {
      const int e = kCells<G>.c[lane + k * G::NT];
      const int cc[3] = {e & 31, (e >> 5) & 31, (e >> 10) & 31};
      cel[k] = cc[0] * G::SX + cc[1] * G::SY + cc[2];
      int g[3];
      bool inner = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        g[a] = o[a] + cc[a] - G::H;
        inner = inner && g[a] >= d.ng && g[a] < d.S[a] - d.ng;
      }
      const int q = (g[0] * d.S[1] + g[1]) * d.S[2] + g[2];
      if (inner) {
        off[k] = (int)boff + q;
        psrc[k] = (int)d.P;
      } else {
        const int2 m = reinterpret_cast<const int2*>(lmap)[q];
        const bool ok = link_entry_ok(m.x, m.y, pool_n);
        off[k] = ok ? m.x : (int)boff + q;
        psrc[k] = ok ? m.y : (int)d.P;
      }
    }
#pragma unroll
    for (int i = 0; i < DEPTH; ++i)
      if (i < ORDER::N) fetch(i, ORDER::at(i));
  }

  // staged slots of field f: the edge bars only for fields with mixed derivatives
  __device__ __forceinline__ static int slots(int f) {
    return needs_edges(f) ? G::PER : G::PER_FACES;
  }

  __device__ __forceinline__ void fetch(int slot, int f) {
    const double* u = F + (long)f * P;
    const int n = slots(f);
#pragma unroll
    for (int k = 0; k < G::PER; ++k) {
      if constexpr (MODE == LINK) {
        // a 32-bit byte offset from the uniform pool base (pools below 2^29 doubles), so the
        // load takes its base from scalar registers and no 64-bit address arithmetic per cell
        const unsigned b = (unsigned)(off[k] + f * psrc[k]) << 3;
        if (k < n) pf[slot][k] = *reinterpret_cast<const double*>(
                                     reinterpret_cast<const char*>(F) + b);
      } else {
        if (k < n) pf[slot][k] = u[off[k]];
      }
    }
    // keep the loads here, ahead of the work they overlap
    __builtin_amdgcn_sched_barrier(0);
  }

  __device__ __forceinline__ void stage(int f) {
    const int at = ORDER::pos(f), slot = DEPTH == 1 ? 0 : at % DEPTH;
    const int n = slots(f);
#pragma unroll
    for (int k = 0; k < G::PER; ++k)
      if (k < n) lds[cel[k]] = pf[slot][k];
    // The wave's LDS accesses complete in issue order; this only stops the compiler from
    // moving this field's reads above the writes.
    __builtin_amdgcn_wave_barrier();
    if (at + DEPTH < ORDER::N) fetch(slot, ORDER::at(at + DEPTH));
  }

  // stage(f) with the slot given at compile time (SLOT must equal ORDER::pos(f) % DEPTH): for
  // per-field loops that are not unrolled, which step DEPTH fields per iteration (a slot index
  // known only at run time would put the in-flight cells in scratch memory)
  template <int SLOT>
  __device__ __forceinline__ void stage_slot(int f) {
    const int at = ORDER::pos(f);
    const int n = slots(f);
#pragma unroll
    for (int k = 0; k < G::PER; ++k)
      if (k < n) lds[cel[k]] = pf[SLOT][k];
    __builtin_amdgcn_wave_barrier();
    if (at + DEPTH < ORDER::N) fetch(SLOT, ORDER::at(at + DEPTH));
  }

  __device__ __forceinline__ double center(int f) const { return Fc[(long)f * P + gc]; }
};

// The ghost points of a padded block, enumerated without the interior: x-ghost planes
// (2 ng x S1 x S2), then x-interior / y-ghost slabs (N0 x 2 ng x S2), then x- and y-interior /
// z-ghost bars (N0 x N1 x 2 ng), N = S - 2 ng; one thread per ghost point keeps every lane of a
// wave busy (a launch over all points would run the ghost work for 3 lanes of the waves at each
// end of every interior z row). Needs an interior of at least one point per axis.
__host__ __device__ inline long ghost_count(const Dims& d) {
  const long g2 = 2L * d.ng, n0 = d.S[0] - g2, n1 = d.S[1] - g2;
  return g2 * d.S[1] * d.S[2] + n0 * g2 * d.S[2] + n0 * n1 * g2;
}

// coordinates of ghost point t < ghost_count(d)
__device__ __forceinline__ void ghost_point(const Dims& d, long t, int* c) {
  const int ng = d.ng, g2 = 2 * ng;
  const long nA = (long)g2 * d.S[1] * d.S[2];
  const long nB = (long)(d.S[0] - g2) * g2 * d.S[2];
  // ghost layer k in [0, 2 ng) of an axis of extent S: the low layers, then the high ones
  auto layer = [ng, g2](int k, int S) { return k < ng ? k : S - g2 + k; };
  if (t < nA) {
    c[2] = (int)(t % d.S[2]);
    t /= d.S[2];
    c[1] = (int)(t % d.S[1]);
    c[0] = layer((int)(t / d.S[1]), d.S[0]);
  } else if (t < nA + nB) {
    t -= nA;
    c[2] = (int)(t % d.S[2]);
    t /= d.S[2];
    c[1] = layer((int)(t % g2), d.S[1]);
    c[0] = ng + (int)(t / g2);
  } else {
    t -= nA + nB;
    const int n1 = d.S[1] - g2;
    c[2] = layer((int)(t % g2), d.S[2]);
    t /= g2;
    c[1] = ng + (int)(t % n1);
    c[0] = ng + (int)(t / n1);
  }
}

// Workgroup -> brick, XCD-aware: the dispatcher deals workgroups round-robin over the 8 XCDs,
// so consecutive ids are spread; remap so that each XCD receives a contiguous run of bricks and
// halo re-reads hit that XCD's L2. Within the run: whole z columns of bricks, the columns in
// tiles of BSSN_BRICK_TILE_XY x BSSN_BRICK_TILE_XY (x fastest), the tiles along x, then y. A
// column's x and y neighbours are then one and T columns away instead of one y row and a whole
// x plane (64 x 64 columns at 256^3, far past the L2): 256^3 RK4 step 38.08 -> 37.38 ms with
// 4 x 4 tiles (2 / 3 / 6 / 8: 37.46 / 37.37 / 37.51 / 37.72 ms); 0 = columns in y, then x.
#ifndef BSSN_BRICK_TILE_XY
#define BSSN_BRICK_TILE_XY 4
#endif

__device__ __forceinline__ bool brick_of_block(int nbx, int nby, int nbz, int* bx, int* by,
                                               int* bz) {
  const int G = nbx * nby * nbz;
  const int G8 = (G + 7) & ~7;
  const int w = (int)blockIdx.x;
  const int L = (w & 7) * (G8 >> 3) + (w >> 3);
  if (L >= G) return false;
  *bz = L % nbz;
  const int t = L / nbz;                   // column index
  if constexpr (BSSN_BRICK_TILE_XY > 0) {
    constexpr int T = BSSN_BRICK_TILE_XY;
    const int ty0 = (t / (nbx * T)) * T;   // first y of this row of tiles, th rows high
    const int th = min(T, nby - ty0);
    const int rr = t - ty0 * nbx;          // column index within the row of tiles
    const int tx0 = (rr / (T * th)) * T;   // first x of this tile, tw columns wide
    const int tw = min(T, nbx - tx0);
    const int q = rr - tx0 * th;           // column index within the tile
    *bx = tx0 + q % tw;
    *by = ty0 + q / tw;
  } else {
    *by = t % nby;
    *bx = t / nby;
  }
  return true;
}

template <class G>
__device__ __forceinline__ bool brick_origin(const Dims& d, int* ox, int* oy, int* oz) {
  const int nbx = (d.n[0] + G::BX - 1) / G::BX;
  const int nby = (d.n[1] + G::BY - 1) / G::BY;
  const int nbz = (d.n[2] + G::BZ - 1) / G::BZ;
  int bx, by, bz;
  if (!brick_of_block(nbx, nby, nbz, &bx, &by, &bz)) return false;
  *ox = d.lo[0] + bx * G::BX;
  *oy = d.lo[1] + by * G::BY;
  *oz = d.lo[2] + bz * G::BZ;
  return true;
}

template <class G>
inline unsigned brick_blocks(const Dims& d) {
  const long nb = (long)((d.n[0] + G::BX - 1) / G::BX) * ((d.n[1] + G::BY - 1) / G::BY) *
                  ((d.n[2] + G::BZ - 1) / G::BZ);
  return (unsigned)((nb + 7) & ~7L);
}

template <class G>
__device__ __forceinline__ bool lane_valid(const Dims& d, int ox, int oy, int oz, int lane) {
  const int tx = G::lx(lane), ty = G::ly(lane), tz = G::lz(lane);
  return (ox + tx < d.lo[0] + d.n[0]) && (oy + ty < d.lo[1] + d.n[1]) &&
         (oz + tz < d.lo[2] + d.n[2]);
}

}  // namespace bssn
