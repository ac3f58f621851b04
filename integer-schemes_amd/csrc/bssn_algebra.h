// The BSSN right-hand side at one grid point, ordered field by field so that each derivative
// is consumed as soon as it is formed and the big intermediates (the Christoffels, the Ricci
// tensor) die early (gfx950 register budget).
//
// Equations: the CAHD + SSL moving-puncture BSSN system the reference evaluates
// (3D/src/bssn3d/vendor/bssneqs_SSL_HD_dxsq.cpp, output statements a_rhs .. B_rhs2, through
// _bssn_rhs_generated.py:bssn_rhs_algebra), written here from the tensor form:
//   Gamma_kij = (d_j gt_ki + d_i gt_kj - d_k gt_ij)/2,  Gamma^m_ij = gt^mk Gamma_kij,
//   CalGt^m = gt^ij Gamma^m_ij,
//   R_ij = Rt_ij + Rchi_ij (conformal Ricci: metric part with the evolved Gt^k in the
//          d_(i Gt_j) term, chi part with CalGt), H = chi gt^ij R_ij - At.At + 2K^2/3,
//   DDalpha_ij = d_i d_j alpha - Gamma3^l_ij d_l alpha (physical Christoffels),
// and the evolution equations for alpha (with slice locking), beta (Gamma-driver), gt, chi (with
// Hamiltonian damping), At, K, Gt, B. Advection is centred. Parity: every output agrees with the
// compiled reference CSE to round-off (tests/test_oracle.py, tests/test_hip_parity.py).
//
// Phase order (fields staged):  gt0..5 | chi | Gt0..2 | alpha | beta0..2 | At0..5 | K | B0..2
//
// Src must provide (all calls uniform across the wave):
//   void   stage(int f)        make field f current
//   double val()               current field at the point
//   double d1(int a)           first derivative of the current field along axis a
//   double d2(int i, int j)    second derivative, i <= j (mixed = d1_j(d1_i))
//   double ko()                Kreiss-Oliger sum over the three axes (KO builds only)
//   double center(int f)       field f at the point, any time
// and may declare `static constexpr bool kEarlyCentres = true` to have the centres of fields
// not staged yet loaded about a phase before their use (a latency-bound source: the LDS
// bricks); otherwise each is loaded where it is used.
// Out must provide: void put(int f, double rhs)   -- called exactly once per field,
//                   void metric_done()             -- after the last gt_ij and At_ij put;
// and may overload out_phase<PH>(Out&), called at the start of phase PH (0 gt, 1 chi, 2 Gt,
// 3 alpha, 4 beta, 5 At, 6 K, 7 B), to issue loads ahead of its puts.
#pragma once

#include <type_traits>

#include "bssn_params.h"

namespace bssn {

// phase hook of bssn_rhs_point: nothing by default
template <int PH, class Out>
__device__ __forceinline__ void out_phase(Out&) {}

template <class S, class = void>
struct EarlyCentres : std::false_type {};
template <class S>
struct EarlyCentres<S, std::void_t<decltype(S::kEarlyCentres)>>
    : std::bool_constant<S::kEarlyCentres> {};

#ifndef BSSN_PIN
#define BSSN_PIN(x) (x) = pin(x)
#endif

__host__ __device__ constexpr int sym(int i, int j) {
  return (i <= j) ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
}
// packed symmetric index -> (i, j)
__host__ __device__ constexpr int pi_(int q) { return q < 3 ? 0 : (q < 5 ? 1 : 2); }
__host__ __device__ constexpr int pj_(int q) { return q < 3 ? q : (q < 5 ? q - 2 : 2); }

// symmetric 3x3 (packed) contraction sum_ij a^ij b_ij
__device__ __forceinline__ double contract_sym(const double* a, const double* b) {
  return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4]);
}

// inverse of a packed symmetric 3x3 matrix; returns the determinant
__device__ __forceinline__ double inverse_sym(const double* g, double* ig) {
  const double a00 = g[3] * g[5] - g[4] * g[4];
  const double a01 = g[2] * g[4] - g[1] * g[5];
  const double a02 = g[1] * g[4] - g[2] * g[3];
  const double a11 = g[0] * g[5] - g[2] * g[2];
  const double a12 = g[1] * g[2] - g[0] * g[4];
  const double a22 = g[0] * g[3] - g[1] * g[1];
  const double det = g[0] * a00 + g[1] * a01 + g[2] * a02;
  const double idet = 1.0 / det;
  ig[0] = a00 * idet; ig[1] = a01 * idet; ig[2] = a02 * idet;
  ig[3] = a11 * idet; ig[4] = a12 * idet; ig[5] = a22 * idet;
  return det;
}

// At^ij (packed), At_ij At^ij and At_ik gt^kl At_lj from At_ij and gt^ij
__device__ __forceinline__ void raise_A(const double* A, const double* ig, double* Aup,
                                        double* AA, double* AkA) {
  double IA[3][3];   // IA[k][j] = gt^kl At_lj
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      IA[k][j] = ig[sym(k, 0)] * A[sym(0, j)] + ig[sym(k, 1)] * A[sym(1, j)] +
                 ig[sym(k, 2)] * A[sym(2, j)];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int i = pi_(q), j = pj_(q);
    Aup[q] = IA[i][0] * ig[sym(0, j)] + IA[i][1] * ig[sym(1, j)] + IA[i][2] * ig[sym(2, j)];
    if (AkA) AkA[q] = A[sym(i, 0)] * IA[0][j] + A[sym(i, 1)] * IA[1][j] + A[sym(i, 2)] * IA[2][j];
  }
  *AA = contract_sym(Aup, A);
}

template <bool KO, class Src, class Out>
__device__ __forceinline__ void bssn_rhs_point(Src& s, const Scal& P, Out& o) {
  // ---------------------------------------------------------------- metric and shift centres
  double g[6], ig[6], b[3];
#pragma unroll
  for (int q = 0; q < 6; ++q) g[q] = s.center(GT0 + q);
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = s.center(BETA0 + i);
  inverse_sym(g, ig);
#pragma unroll
  for (int q = 0; q < 6; ++q) BSSN_PIN(ig[q]);

  // ---------------------------------------------------------------- gt_ij: Christoffels, Rt
  out_phase<0>(o);
  double Rt[6];        // accumulates the Ricci tensor
  double rg[6];        // gt right-hand sides (advection + KO so far)
  double C2[3][6], CalGt[3];
  // Early centres: the centre values of fields not staged yet are loaded about a phase before
  // their use, so their latency hides behind work (LDS bricks at 256^3 FD4/KO6: RK4 step
  // 40.7 -> 40.0 ms, bare RHS 6.36 -> 6.12 ms). For the bundle sources, whose derivative loads
  // already keep the memory system busy, the early loads only add live registers (two-pass RHS
  // 19.3 -> 21.8 ms), so they load each centre where it is used.
  constexpr bool EC = EarlyCentres<Src>::value;
  double Acen[6];      // At_ij: with the fourth gt field (early) or before the Christoffels die
  {
    double dg[3][6];   // d_k gt_q
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      s.stage(GT0 + q);
      if (EC && q == 3) {
#pragma unroll
        for (int p = 0; p < 6; ++p) Acen[p] = s.center(AT0 + p);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) dg[a][q] = s.d1(a);
      const double lapg = ig[0] * s.d2(0, 0) + ig[3] * s.d2(1, 1) + ig[5] * s.d2(2, 2) +
                          2.0 * (ig[1] * s.d2(0, 1) + ig[2] * s.d2(0, 2) + ig[4] * s.d2(1, 2));
      Rt[q] = -0.5 * lapg;
      double r = b[0] * dg[0][q] + b[1] * dg[1][q] + b[2] * dg[2][q];
      if (KO) r += s.ko();
      rg[q] = pin(r);
    }
    double C1[3][6];   // Gamma_k(ij), first index lowered
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int i = pi_(q), j = pj_(q);
        C1[k][q] = 0.5 * (dg[j][sym(k, i)] + dg[i][sym(k, j)] - dg[k][q]);
      }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 6; ++q)
        C2[m][q] = ig[sym(m, 0)] * C1[0][q] + ig[sym(m, 1)] * C1[1][q] + ig[sym(m, 2)] * C1[2][q];
#pragma unroll
    for (int m = 0; m < 3; ++m) CalGt[m] = contract_sym(ig, C2[m]);
    //   Rt_ij += CalGt^k (Gamma_ijk + Gamma_jik)/2
    //          + gt^lm (Gamma^k_li Gamma_jkm + Gamma^k_lj Gamma_ikm + Gamma^k_im Gamma_klj)
    // with the raised forms Gu[i][k][m] = gt^ml Gamma^k_li and gt^ml Gamma_klj.
    double Gu[3][3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m)
          Gu[i][k][m] = ig[sym(m, 0)] * C2[k][sym(0, i)] + ig[sym(m, 1)] * C2[k][sym(1, i)] +
                        ig[sym(m, 2)] * C2[k][sym(2, i)];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) t += CalGt[k] * (C1[i][sym(j, k)] + C1[j][sym(i, k)]);
      double u = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const double gd = ig[sym(m, 0)] * C1[k][sym(0, j)] + ig[sym(m, 1)] * C1[k][sym(1, j)] +
                            ig[sym(m, 2)] * C1[k][sym(2, j)];
          u += Gu[i][k][m] * C1[j][sym(k, m)] + Gu[j][k][m] * C1[i][sym(k, m)] +
               C2[k][sym(i, m)] * gd;
        }
      Rt[q] += 0.5 * t + u;
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    BSSN_PIN(Rt[q]);
    BSSN_PIN(C2[0][q]); BSSN_PIN(C2[1][q]); BSSN_PIN(C2[2][q]);
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) BSSN_PIN(CalGt[m]);

  // At^ij from the At centres, for the two contractions needed before the Christoffels die:
  // Gamma^i_jk At^jk and At_ij At^ij.
  double CA[3], AA;
  {
    if (!EC) {
#pragma unroll
      for (int p = 0; p < 6; ++p) Acen[p] = s.center(AT0 + p);
    }
    double Aup[6];
    raise_A(Acen, ig, Aup, &AA, nullptr);
#pragma unroll
    for (int i = 0; i < 3; ++i) CA[i] = pin(contract_sym(C2[i], Aup));
    BSSN_PIN(AA);
  }

  // ---------------------------------------------------------------- chi: Ricci chi part
  out_phase<1>(o);
  double chi, dchi[3];
  double rchi;
  {
    s.stage(CHI);
    chi = s.val();
#pragma unroll
    for (int a = 0; a < 3; ++a) dchi[a] = s.d1(a);
    double igdchi[3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
      igdchi[m] = ig[sym(m, 0)] * dchi[0] + ig[sym(m, 1)] * dchi[1] + ig[sym(m, 2)] * dchi[2];
    double d2c[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) d2c[q] = s.d2(pi_(q), pj_(q));
    const double ichi = 1.0 / chi;
    const double dcdc = dchi[0] * igdchi[0] + dchi[1] * igdchi[1] + dchi[2] * igdchi[2];
    const double lapchi = contract_sym(ig, d2c) - 1.5 * ichi * dcdc -
                          (CalGt[0] * dchi[0] + CalGt[1] * dchi[1] + CalGt[2] * dchi[2]);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      const double c = C2[0][q] * dchi[0] + C2[1][q] * dchi[1] + C2[2][q] * dchi[2];
      Rt[q] += 0.5 * ichi * (d2c[q] - c) - 0.25 * ichi * ichi * dchi[i] * dchi[j] +
               0.5 * ichi * g[q] * lapchi;
    }
    rchi = b[0] * dchi[0] + b[1] * dchi[1] + b[2] * dchi[2];
    if (KO) rchi += s.ko();
#pragma unroll
    for (int q = 0; q < 6; ++q) BSSN_PIN(Rt[q]);
    BSSN_PIN(rchi);
  }

  // ---------------------------------------------------------------- Gt^k: d_(i Gt_j) term
  out_phase<2>(o);
  double K = EC ? s.center(KTR) : 0.0;
  double advGt[3], koGt[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.stage(GTI0 + k);
    double dG[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dG[a] = s.d1(a);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      Rt[q] += 0.5 * (g[sym(k, i)] * dG[j] + g[sym(k, j)] * dG[i]);
    }
    advGt[k] = pin(b[0] * dG[0] + b[1] * dG[1] + b[2] * dG[2]);
    koGt[k] = KO ? s.ko() : 0.0;
  }
  // R_ij is complete from here on.

  // ---------------------------------------------------------------- alpha: D_iD_j alpha
  out_phase<3>(o);
  if (!EC) K = s.center(KTR);
  double Bcen[3];      // B^i: here (early) or in the beta phase
  if (EC) {
#pragma unroll
    for (int i = 0; i < 3; ++i) Bcen[i] = s.center(BB0 + i);
  }
  double al, dal[3], rGt[3], rK, rA[6];
  {
    s.stage(ALPHA);
    al = s.val();
#pragma unroll
    for (int a = 0; a < 3; ++a) dal[a] = s.d1(a);
    const double ichi = 1.0 / chi;
    double igdd = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      igdd += (ig[sym(m, 0)] * dchi[0] + ig[sym(m, 1)] * dchi[1] + ig[sym(m, 2)] * dchi[2]) *
              dal[m];
    double DD[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      const double c3 = C2[0][q] * dal[0] + C2[1][q] * dal[1] + C2[2][q] * dal[2] -
                        0.5 * ichi * (dal[i] * dchi[j] + dal[j] * dchi[i] - g[q] * igdd);
      DD[q] = s.d2(i, j) - c3;
    }
    const double Rs = contract_sym(ig, Rt);
    double X[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) X[q] = al * Rt[q] - DD[q];
    const double trX3 = contract_sym(ig, X) * (1.0 / 3.0);
#pragma unroll
    for (int q = 0; q < 6; ++q) rA[q] = pin(chi * (X[q] - g[q] * trX3));
    rK = pin(-chi * contract_sym(ig, DD));
    // Hamiltonian constraint and its damping of chi (CAHD)
    const double ham = chi * Rs - AA + (2.0 / 3.0) * K * K;
    rchi += P.cahd * chi * ham;
    // lapse
    const double sq = sqrt(chi);
    double ra = P.lam[0] * (b[0] * dal[0] + b[1] * dal[1] + b[2] * dal[2]) - 2.0 * al * K -
                sq * P.ssl_h * (al - sq) * P.ssl_e;
    if (KO) ra += s.ko();
    o.put(ALPHA, ra);
#pragma unroll
    for (int i = 0; i < 3; ++i) rGt[i] = pin(advGt[i] + 2.0 * al * CA[i]);
  }
  // The Christoffels and the Ricci tensor are dead from here on.

  // ---------------------------------------------------------------- beta^i
  out_phase<4>(o);
  double db[3][3];   // db[j][i] = d_j beta^i
  {
    const double Bf = 0.75 * (P.lamf[0] + P.lamf[1] * al);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      s.stage(BETA0 + i);
#pragma unroll
      for (int a = 0; a < 3; ++a) db[a][i] = s.d1(a);
      double d2b[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) d2b[q] = s.d2(pi_(q), pj_(q));
      rGt[i] += contract_sym(ig, d2b);
      // (1/3) gt^mj d_j d_i beta^i
#pragma unroll
      for (int m = 0; m < 3; ++m)
        rGt[m] += (1.0 / 3.0) * (ig[sym(m, 0)] * d2b[sym(0, i)] + ig[sym(m, 1)] * d2b[sym(1, i)] +
                                 ig[sym(m, 2)] * d2b[sym(2, i)]);
      if (!EC) Bcen[i] = s.center(BB0 + i);
      double rb = Bf * Bcen[i] +
                  P.lam[1] * (b[0] * db[0][i] + b[1] * db[1][i] + b[2] * db[2][i]);
      if (KO) rb += s.ko();
      o.put(BETA0 + i, rb);
    }
  }
  const double divb = db[0][0] + db[1][1] + db[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    rGt[i] += -(CalGt[0] * db[0][i] + CalGt[1] * db[1][i] + CalGt[2] * db[2][i]) +
              (2.0 / 3.0) * CalGt[i] * divb;
  rchi += (2.0 / 3.0) * chi * (al * K - divb);
  o.put(CHI, rchi);

  // ---------------------------------------------------------------- At_ij, and gt_ij
  out_phase<5>(o);
  {
    double A[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      s.stage(AT0 + q);
      A[q] = s.val();
      double r = b[0] * s.d1(0) + b[1] * s.d1(1) + b[2] * s.d1(2);
      if (KO) r += s.ko();
      rA[q] += r;
    }
    double Aup[6], AA2, AkA[6];
    raise_A(A, ig, Aup, &AA2, AkA);
    const double ichi = 1.0 / chi;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double t = 0.0, w = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        t += Aup[sym(i, j)] * dal[j];
        w += Aup[sym(i, j)] * dchi[j];
      }
      rGt[i] += -2.0 * t - 3.0 * al * ichi * w;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      double lie = 0.0, lieg = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lie += A[sym(i, k)] * db[j][k] + A[sym(k, j)] * db[i][k];
        lieg += g[sym(i, k)] * db[j][k] + g[sym(k, j)] * db[i][k];
      }
      o.put(AT0 + q, rA[q] + lie - (2.0 / 3.0) * A[q] * divb + al * (K * A[q] - 2.0 * AkA[q]));
      o.put(GT0 + q, rg[q] + lieg - (2.0 / 3.0) * g[q] * divb - 2.0 * al * A[q]);
    }
    o.metric_done();
  }

  // ---------------------------------------------------------------- K
  out_phase<6>(o);
  {
    s.stage(KTR);
    double dK[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dK[a] = s.d1(a);
    double r = rK + b[0] * dK[0] + b[1] * dK[1] + b[2] * dK[2] + al * (K * K * (1.0 / 3.0) + AA);
    if (KO) r += s.ko();
    o.put(KTR, r);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      rGt[i] -= (4.0 / 3.0) * al *
                (ig[sym(i, 0)] * dK[0] + ig[sym(i, 1)] * dK[1] + ig[sym(i, 2)] * dK[2]);
  }

  // ---------------------------------------------------------------- B^i and Gt^i
  out_phase<7>(o);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    s.stage(BB0 + i);
    double r = rGt[i] - P.eta * s.val() +
               P.lam[2] * (b[0] * s.d1(0) + b[1] * s.d1(1) + b[2] * s.d1(2)) - P.lam[3] * advGt[i];
    if (KO) r += s.ko();
    o.put(BB0 + i, r);
    o.put(GTI0 + i, rGt[i] + koGt[i]);
  }
}

// Hamiltonian and momentum constraints at one point (3D/src/bssn3d/constraints.py:38-51, the
// reference's physcon CSE 3D/src/bssn3d/vendor/physcon.cpp:360-1287):
//   H   = chi gt^ij R_ij - At_ij At^ij + 2K^2/3
//   M^i = gt^jk (d_j At_ki - Gamma^m_ji At_km) - Gt^j At_ij - 3/(2chi) gt^jk At_ki d_j chi
//         - 2/3 d_i K
// Same Src interface as bssn_rhs_point (no KO); Out receives put(0..3, value).
template <class Src, class Out>
__device__ __forceinline__ void bssn_constraints_point(Src& s, Out& o) {
  double g[6], ig[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) g[q] = s.center(GT0 + q);
  {
    const double a00 = g[3] * g[5] - g[4] * g[4];
    const double a01 = g[2] * g[4] - g[1] * g[5];
    const double a02 = g[1] * g[4] - g[2] * g[3];
    const double a11 = g[0] * g[5] - g[2] * g[2];
    const double a12 = g[1] * g[2] - g[0] * g[4];
    const double a22 = g[0] * g[3] - g[1] * g[1];
    const double idet = 1.0 / (g[0] * a00 + g[1] * a01 + g[2] * a02);
    ig[0] = a00 * idet; ig[1] = a01 * idet; ig[2] = a02 * idet;
    ig[3] = a11 * idet; ig[4] = a12 * idet; ig[5] = a22 * idet;
  }
  double dg[3][6], Rt[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    s.stage(GT0 + q);
#pragma unroll
    for (int a = 0; a < 3; ++a) dg[a][q] = s.d1(a);
    Rt[q] = -0.5 * (ig[0] * s.d2(0, 0) + ig[3] * s.d2(1, 1) + ig[5] * s.d2(2, 2) +
                    2.0 * (ig[1] * s.d2(0, 1) + ig[2] * s.d2(0, 2) + ig[4] * s.d2(1, 2)));
  }
  double C2[3][6], CalGt[3];
  {
    double C1[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int i = pi_(q), j = pj_(q);
        C1[k][q] = 0.5 * (dg[j][sym(k, i)] + dg[i][sym(k, j)] - dg[k][q]);
      }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 6; ++q)
        C2[m][q] = ig[sym(m, 0)] * C1[0][q] + ig[sym(m, 1)] * C1[1][q] + ig[sym(m, 2)] * C1[2][q];
#pragma unroll
    for (int m = 0; m < 3; ++m) CalGt[m] = contract_sym(ig, C2[m]);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      double t = 0.0, u = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) t += CalGt[k] * (C1[i][sym(j, k)] + C1[j][sym(i, k)]);
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const double gui = ig[sym(m, 0)] * C2[k][sym(0, i)] + ig[sym(m, 1)] * C2[k][sym(1, i)] +
                             ig[sym(m, 2)] * C2[k][sym(2, i)];
          const double guj = ig[sym(m, 0)] * C2[k][sym(0, j)] + ig[sym(m, 1)] * C2[k][sym(1, j)] +
                             ig[sym(m, 2)] * C2[k][sym(2, j)];
          const double gd = ig[sym(m, 0)] * C1[k][sym(0, j)] + ig[sym(m, 1)] * C1[k][sym(1, j)] +
                            ig[sym(m, 2)] * C1[k][sym(2, j)];
          u += gui * C1[j][sym(k, m)] + guj * C1[i][sym(k, m)] + C2[k][sym(i, m)] * gd;
        }
      Rt[q] += 0.5 * t + u;
    }
  }
  double chi, dchi[3];
  {
    s.stage(CHI);
    chi = s.val();
#pragma unroll
    for (int a = 0; a < 3; ++a) dchi[a] = s.d1(a);
    double d2c[6], igdchi[3];
#pragma unroll
    for (int q = 0; q < 6; ++q) d2c[q] = s.d2(pi_(q), pj_(q));
#pragma unroll
    for (int m = 0; m < 3; ++m)
      igdchi[m] = ig[sym(m, 0)] * dchi[0] + ig[sym(m, 1)] * dchi[1] + ig[sym(m, 2)] * dchi[2];
    const double ichi = 1.0 / chi;
    const double lapchi = contract_sym(ig, d2c) -
                          1.5 * ichi * (dchi[0] * igdchi[0] + dchi[1] * igdchi[1] + dchi[2] * igdchi[2]) -
                          (CalGt[0] * dchi[0] + CalGt[1] * dchi[1] + CalGt[2] * dchi[2]);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      const double c = C2[0][q] * dchi[0] + C2[1][q] * dchi[1] + C2[2][q] * dchi[2];
      Rt[q] += 0.5 * ichi * (d2c[q] - c) - 0.25 * ichi * ichi * dchi[i] * dchi[j] +
               0.5 * ichi * g[q] * lapchi;
    }
  }
  double Gt[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.stage(GTI0 + k);
    Gt[k] = s.val();
    double dG[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dG[a] = s.d1(a);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      Rt[q] += 0.5 * (g[sym(k, i)] * dG[j] + g[sym(k, j)] * dG[i]);
    }
  }
  double A[6], mom[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    s.stage(AT0 + q);
    A[q] = s.val();
    double dA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dA[a] = s.d1(a);
    // gt^jk d_j At_ki for every ordered (k, i) with At_ki = At_q
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if (sym(k, i) == q)
          mom[i] += ig[sym(0, k)] * dA[0] + ig[sym(1, k)] * dA[1] + ig[sym(2, k)] * dA[2];
  }
  double AA;
  {
    double IA[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        IA[k][j] = ig[sym(k, 0)] * A[sym(0, j)] + ig[sym(k, 1)] * A[sym(1, j)] +
                   ig[sym(k, 2)] * A[sym(2, j)];
    double Aup[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int i = pi_(q), j = pj_(q);
      Aup[q] = IA[i][0] * ig[sym(0, j)] + IA[i][1] * ig[sym(1, j)] + IA[i][2] * ig[sym(2, j)];
    }
    AA = contract_sym(Aup, A);
    const double ichi = 1.0 / chi;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double c = 0.0, gA = 0.0, w = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double cm = C2[0][sym(j, i)] * A[sym(k, 0)] + C2[1][sym(j, i)] * A[sym(k, 1)] +
                            C2[2][sym(j, i)] * A[sym(k, 2)];
          c += ig[sym(j, k)] * cm;
          w += ig[sym(j, k)] * A[sym(k, i)] * dchi[j];
        }
#pragma unroll
      for (int j = 0; j < 3; ++j) gA += Gt[j] * A[sym(i, j)];
      mom[i] += -c - gA - 1.5 * ichi * w;
    }
  }
  {
    s.stage(KTR);
    const double K = s.val();
#pragma unroll
    for (int i = 0; i < 3; ++i) mom[i] -= (2.0 / 3.0) * s.d1(i);
    const double ham = chi * contract_sym(ig, Rt) - AA + (2.0 / 3.0) * K * K;
    o.put(0, ham);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) o.put(1 + i, mom[i]);
}

}  // namespace bssn
