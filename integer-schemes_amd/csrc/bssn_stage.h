// Stage write-back of the fused kernels: the bare RHS, the RK4 and MSRK stage updates, and
// the algebraic enforcement folded into a step's final store.
//   RK4 stages  : BSSNEvolution.step (3D/src/bssn3d/evolve.py:125-132)
//   MSRK stages : BSSNEvolution._msrk2_step (evolve.py:139-153)
//   enforce     : BSSNEvolution.enforce (evolve.py:93-119)
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

#include <type_traits>

#include "bssn_kernels.h"
#include "bssn_params.h"

namespace bssn {

__device__ __forceinline__ void enforce_point(double* g, double* a) {
  // det(gt) = 1 and tr(At) = 0 (3D/src/bssn3d/evolve.py:93-119)
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
  for (int i = 0; i < 6; ++i) a[i] = a[i] - g[i] * third;
}

// Waves per SIMD the register allocator targets: 1 gives a wave the whole 512-entry
// VGPR+AGPR file (no spills, less latency hiding), 2 halves it. Measured at 256^3 (FD4/KO6):
// the RK stages (KO + stage update, write-back operands preloaded) are faster at 1
// (43.0 vs 51.7 ms per step), the bare RHS slightly (7.09 vs 7.27 ms).
#ifndef BSSN_RHS_WAVES
#define BSSN_RHS_WAVES 1
#endif
#ifndef BSSN_STAGE_WAVES
#define BSSN_STAGE_WAVES 1
#endif

template <int KIND>
constexpr int fused_waves() {
  return KIND == KIND_RHS ? BSSN_RHS_WAVES : BSSN_STAGE_WAVES;
}

// Fields in flight while one is computed (at 1 wave per SIMD, where no other wave hides the
// load latency). Measured at 256^3 (FD4/KO6): the bare RHS is faster with two fields in flight
// (6.37 vs 7.08 ms); the RK stages, which also hold write-back operands, spill with two and
// are slower (42.5 vs 40.7 ms per step), so they keep one.
#ifndef BSSN_RHS_DEPTH
#define BSSN_RHS_DEPTH 2
#endif
#ifndef BSSN_STAGE_DEPTH
#define BSSN_STAGE_DEPTH 1
#endif
// FD order 8 (19 staged cells per lane) keeps one field in flight in every kind: its stencils
// alone push the stage kinds past the register file.
template <int KIND, int ORD = 4>
constexpr int fused_depth() {
  return fused_waves<KIND>() != 1 || ORD == 8
             ? 1
             : (KIND == KIND_RHS ? BSSN_RHS_DEPTH : BSSN_STAGE_DEPTH);
}

// Write-back operands are loaded this many phases of bssn_rhs_point before the phase that
// stores their field (phases: 0 gt, 1 chi, 2 Gt, 3 alpha, 4 beta, 5 At, 6 K, 7 B). Measured at
// 256^3 (FD4/KO6), ms per RK4 / MSRK step: all at kernel start 43.0 / 35.1, lead 0 40.9 / 32.7,
// lead 1 40.8 / 32.7, lead 2 40.7 / 32.6: loaded at the start they stay live through the
// register-heaviest phases (the Christoffels and the Ricci tensor) and push values into AGPRs.
#ifndef BSSN_WB_LEAD
#define BSSN_WB_LEAD 2
#endif

// The write-back kind is a template parameter: each launch runs the code of its own kind only
// (a runtime switch over all kinds made the stage kernel half again as large).
// RK2 and RK3 share their code (c0 differs); KIND_RK3 launches StageOut<KIND_RK2>.
template <int KIND>
struct StageOut {
  const StageArgs& A;
  long P, gp;
  bool valid;
  double gnew[6], anew[6];

  // The write-back operands are loaded ahead of their stores, so their latency is hidden
  // behind the evaluation instead of stalling each put: all at construction (deferred =
  // false), or, under bssn_rhs_point, field by field BSSN_WB_LEAD phases before the phase
  // that stores the field (deferred = true, loads issued by phase<PH>()).
  static constexpr bool NU = KIND != KIND_RHS;
  static constexpr bool NK = KIND == KIND_RK2 || KIND == KIND_RK4 || KIND == KIND_MS2;
  static constexpr bool NF = KIND == KIND_MS1 || KIND == KIND_MS2 || KIND == KIND_MS3;
  double pu[NU ? BSSN_NF : 1], pk[NK ? BSSN_NF : 1], pp[NF ? BSSN_NF : 1];
  bool deferred;

  __device__ __forceinline__ StageOut(const StageArgs& a, long P_, long gp_, bool v,
                                      bool defer = false)
      : A(a), P(P_), gp(gp_), valid(v), deferred(defer) {
#pragma unroll
    for (int f = 0; f < BSSN_NF; ++f)
      if (!defer) load(f);
  }

  __device__ __forceinline__ void load(int f) {
    const long i = (long)f * P + gp;
    if constexpr (NU) pu[f] = A.U[i];
    if constexpr (NK) pk[f] = A.K1[i];
    if constexpr (NF) pp[f] = A.FP[i];
  }

  // the bssn_rhs_point phase that stores field f, and the one that loads its operands
  static constexpr int put_phase(int f) {
    return f == ALPHA ? 3
         : (f >= BETA0 && f <= BETA2) || f == CHI ? 4
         : (f >= GT0 && f <= GT5) || (f >= AT0 && f <= AT5) ? 5
         : f == KTR ? 6
         : 7;   // B^i, Gt^i
  }
  static constexpr int load_phase(int f) {
    return put_phase(f) - BSSN_WB_LEAD < 0 ? 0 : put_phase(f) - BSSN_WB_LEAD;
  }

  // called by bssn_rhs_point at the start of phase PH
  template <int PH>
  __device__ __forceinline__ void phase() {
    if (!deferred) return;
#pragma unroll
    for (int f = 0; f < BSSN_NF; ++f)
      if (load_phase(f) == PH) load(f);
  }

  __device__ __forceinline__ void final_value(int f, double u) {
    if (f >= GT0 && f <= GT5) {
      gnew[f - GT0] = u;
    } else if (f >= AT0 && f <= AT5) {
      anew[f - AT0] = u;
    } else {
      if (f == CHI) u = u < A.chi_floor ? A.chi_floor : u;
      if (f == ALPHA) u = u < A.alpha_floor ? A.alpha_floor : u;
      A.Uout[(long)f * P + gp] = u;
    }
  }

  __device__ __forceinline__ void put(int f, double k) {
    if (!valid) return;
    const long i = (long)f * P + gp;
    if constexpr (KIND == KIND_RHS) {
      A.Y[i] = k;
    } else if constexpr (KIND == KIND_RK1) {            // y + (dt/2) k1 ; S = k1
      A.Y[i] = pu[f] + A.c0 * k;
      A.K1[i] = k;
    } else if constexpr (KIND == KIND_RK2) {            // y + c0 k ; S = S + 2 k  (RK2, RK3)
      A.Y[i] = pu[f] + A.c0 * k;
      A.K1[i] = pk[f] + 2.0 * k;
    } else if constexpr (KIND == KIND_RK4) {            // y + (dt/6) (S + k4)
      final_value(f, pu[f] + A.c0 * (pk[f] + k));
    } else if constexpr (KIND == KIND_MS1) {            // y + dt (a20 k0 + a21 k1) ; store k1
      A.Y[i] = pu[f] + A.dt * (A.c0 * pp[f] + A.c1 * k);
      A.K1[i] = k;
    } else if constexpr (KIND == KIND_MS2) {            // y + dt (a30 k0 + a31 k1 + a32 k2)
      const double k0 = pp[f], k1 = pk[f];
      A.Y[i] = pu[f] + A.dt * (A.c0 * k0 + A.c1 * k1 + A.c2 * k);
      A.FP[i] = MS_B0 * k0 + MS_B1 * k1 + MS_B2 * k;
    } else {                                            // MS3: y + dt (b0 k0 + ... + b3 k3)
      static_assert(KIND == KIND_MS3, "unknown stage kind");
      final_value(f, pu[f] + A.dt * (pp[f] + MS_B3 * k));
    }
  }

  // all gt_ij and At_ij of a final stage are known: enforce det = 1, tr = 0 and store them
  __device__ __forceinline__ void metric_done() {
    if constexpr (KIND == KIND_RK4 || KIND == KIND_MS3) {
      if (!valid) return;
      enforce_point(gnew, anew);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        A.Uout[(long)(GT0 + q) * P + gp] = gnew[q];
        A.Uout[(long)(AT0 + q) * P + gp] = anew[q];
      }
    }
  }
};

// bssn_rhs_point's phase hook for the stage write-back (more specialised than its generic,
// empty hook)
template <int PH, int KIND>
__device__ __forceinline__ void out_phase(StageOut<KIND>& o) {
  o.template phase<PH>();
}

// Call fn(std::integral_constant<int, K>) with K the template kind of the runtime stage kind
// (MSRK kinds only if MSRK is set).
template <bool MSRK, class Fn>
__host__ inline hipError_t with_kind(int kind, Fn&& fn) {
  switch (kind) {
    case KIND_RHS: return fn(std::integral_constant<int, KIND_RHS>());
    case KIND_RK1: return fn(std::integral_constant<int, KIND_RK1>());
    case KIND_RK2:
    case KIND_RK3: return fn(std::integral_constant<int, KIND_RK2>());
    case KIND_RK4: return fn(std::integral_constant<int, KIND_RK4>());
    default: break;
  }
  if constexpr (MSRK) {
    switch (kind) {
      case KIND_MS1: return fn(std::integral_constant<int, KIND_MS1>());
      case KIND_MS2: return fn(std::integral_constant<int, KIND_MS2>());
      case KIND_MS3: return fn(std::integral_constant<int, KIND_MS3>());
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace bssn
