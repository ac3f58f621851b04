/*
 * CPU oracle for the BSSN right-hand side and its time step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's verbatim path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, only ever as the checker.
 * The product (integer-schemes_amd/) never links, loads or calls anything here.
 *
 * What it restates (reference paths relative to the reference root):
 *   - centred FD stencils with edge padding       common/src/mcs_common/derivatives.py:33-106
 *   - the 138-derivative bundle, mixed = d1_j d1_i 3D/src/bssn3d/derivative_bundle.py:31-48
 *   - the CAHD+SSL BSSN algebra                    3D/src/bssn3d/vendor/bssneqs_SSL_HD_dxsq.cpp
 *     (the Dendro-GR CSE the reference transliterates in _bssn_rhs_generated.py), written here
 *     from the tensor equations (conformal Christoffels, conformal Ricci split into the metric
 *     and chi parts, trace-free projection, Gamma-driver shift), not from the CSE text
 *   - Kreiss-Oliger dissipation                    3D/src/bssn3d/evolve.py:76-90
 *   - periodic ghost sync                          3D/src/bssn3d/evolve.py:65-73
 *   - algebraic enforcement                        3D/src/bssn3d/evolve.py:93-119
 *   - classic RK4 step                             3D/src/bssn3d/evolve.py:125-132
 *   - MSRK rk4_2_2 step                            3D/src/bssn3d/evolve.py:139-153,
 *                                                  2D/src/mcs2d/msrk.py:47-52
 *   - Hamiltonian/momentum constraints             3D/src/bssn3d/constraints.py:38-62,
 *                                                  3D/src/bssn3d/vendor/physcon.cpp:360-1287
 *
 * Parity pins (tests/test_oracle_*.py): the algebra and the constraints against the
 * reference's compiled Dendro CSE on the reference oracle's own seeded inputs
 * (3D/src/bssn3d/oracle.py:46-85; golden vectors in tests/golden/), Minkowski staticity,
 * stencil coefficients and polynomial exactness, MSRK order conditions and RK4 order.
 */
#include "bssn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_LONG
#define SQRT sqrtl
#define EXP expl
#define POW powl
#else
#define SQRT sqrt
#define EXP exp
#define POW pow
#endif

#define NF 24
#define ND 138

/* field indices (3D/src/bssn3d/state.py:26-34) */
enum { ALPHA = 0, CHI = 1, KK = 2, GT0 = 3, BETA0 = 9, AT0 = 12, GTI0 = 18, BB0 = 21 };

/* fields that carry second derivatives, in state order */
static const int S2F[11] = {ALPHA, CHI, GT0, GT0 + 1, GT0 + 2, GT0 + 3, GT0 + 4, GT0 + 5,
                            BETA0, BETA0 + 1, BETA0 + 2};
/* symmetric pair -> packed index (state.py:52-55) */
static const int SYM[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
static const int PAIR_I[6] = {0, 0, 0, 1, 1, 2};
static const int PAIR_J[6] = {0, 1, 2, 1, 2, 2};

int orc_version(void) { return 1; }

/* OpenMP team size of the parallel loops below (timing runs on 1 core and on all cores). */
#ifdef _OPENMP
#include <omp.h>
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
#else
void orc_set_threads(int n) { (void)n; }
int orc_get_threads(void) { return 1; }
#endif

/* ---------------------------------------------------------------- stencils */

int orc_stencil(int kind, int order, real* c) {
  static const real c1_4[5] = {1, -8, 0, 8, -1};
  static const real c1_6[7] = {-1, 9, -45, 0, 45, -9, 1};
  static const real c1_8[9] = {3, -32, 168, -672, 0, 672, -168, 32, -3};
  static const real c2_4[5] = {-1, 16, -30, 16, -1};
  static const real c2_6[7] = {2, -27, 270, -490, 270, -27, 2};
  static const real c2_8[9] = {-9, 128, -1008, 8064, -14350, 8064, -1008, 128, -9};
  static const real ko_4[5] = {-1, 4, -6, 4, -1};
  static const real ko_6[7] = {1, -6, 15, -20, 15, -6, 1};
  static const real ko_8[9] = {-1, 8, -28, 56, -70, 56, -28, 8, -1};
  const real* src = NULL;
  real den = 1.0;
  if (kind == 0) {
    if (order == 4) { src = c1_4; den = 12.0; }
    if (order == 6) { src = c1_6; den = 60.0; }
    if (order == 8) { src = c1_8; den = 840.0; }
  } else if (kind == 1) {
    if (order == 4) { src = c2_4; den = 12.0; }
    if (order == 6) { src = c2_6; den = 180.0; }
    if (order == 8) { src = c2_8; den = 5040.0; }
  } else if (kind == 2) {
    if (order == 4) { src = ko_4; den = 16.0; }
    if (order == 6) { src = ko_6; den = 64.0; }
    if (order == 8) { src = ko_8; den = 256.0; }
  }
  if (!src) return -1;
  for (int k = 0; k <= order; ++k) c[k] = src[k] / den;
  return order + 1;
}

static inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

typedef struct {
  int S[3];
  long P;
} Dims;

static inline long lin(const Dims* d, int x, int y, int z) {
  return ((long)x * d->S[1] + y) * d->S[2] + z;
}

/* value of u at p + off*e_axis, clamped along that axis */
static inline real tap(const real* u, const Dims* d, const int p[3], int axis, int off) {
  int q[3] = {p[0], p[1], p[2]};
  q[axis] = clampi(q[axis] + off, d->S[axis] - 1);
  return u[lin(d, q[0], q[1], q[2])];
}

/* one-axis stencil at one point: (coeffs * factor) first, then a left-to-right sum that
 * includes every tap (derivatives.py:79,87-88) */
static real stencil_at(const real* u, const Dims* d, const int p[3], int axis,
                         const real* w, int ntap) {
  const int r = (ntap - 1) / 2;
  real s = 0.0;
  for (int k = 0; k < ntap; ++k) s += w[k] * tap(u, d, p, axis, k - r);
  return s;
}

/* mixed derivative d1_j(d1_i u) with i < j, each application edge-padded on its own axis */
static real mixed_at(const real* u, const Dims* d, const int p[3], int i, int j,
                       const real* wi, const real* wj, int ntap) {
  const int r = (ntap - 1) / 2;
  real s = 0.0;
  for (int b = 0; b < ntap; ++b) {
    int q[3] = {p[0], p[1], p[2]};
    q[j] = clampi(q[j] + b - r, d->S[j] - 1);
    s += wj[b] * stencil_at(u, d, q, i, wi, ntap);
  }
  return s;
}

void orc_apply_axis(const real* u, real* out, int Sx, int Sy, int Sz, int axis, int kind,
                    int order, real factor) {
  real c[9], w[9];
  const int nt = orc_stencil(kind, order, c);
  if (nt < 0) return;
  for (int k = 0; k < nt; ++k) w[k] = c[k] * factor;
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < Sx; ++x)
    for (int y = 0; y < Sy; ++y)
      for (int z = 0; z < Sz; ++z) {
        const int p[3] = {x, y, z};
        out[lin(&d, x, y, z)] = stencil_at(u, &d, p, axis, w, nt);
      }
}

typedef struct {
  int nt;
  real w1[3][9]; /* first-derivative weights per axis (C1 / h) */
  real w2[3][9]; /* second-derivative weights per axis (C2 / h^2) */
} FDW;

static int fd_weights(FDW* fw, const real* h3, int order) {
  real c1[9], c2[9];
  fw->nt = orc_stencil(0, order, c1);
  if (fw->nt < 0 || orc_stencil(1, order, c2) < 0) return -1;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < fw->nt; ++k) {
      fw->w1[a][k] = c1[k] * (1.0 / h3[a]);
      fw->w2[a][k] = c2[k] * (1.0 / (h3[a] * h3[a]));
    }
  return 0;
}

/* all 138 derivatives at one point; Dp[d * stride] */
static void point_derivs(const real* F, const Dims* d, const FDW* fw, const int p[3],
                         real* Dp, long stride) {
  for (int f = 0; f < NF; ++f) {
    const real* u = F + (long)f * d->P;
    for (int a = 0; a < 3; ++a)
      Dp[(3 * f + a) * stride] = stencil_at(u, d, p, a, fw->w1[a], fw->nt);
  }
  for (int s = 0; s < 11; ++s) {
    const real* u = F + (long)S2F[s] * d->P;
    for (int q = 0; q < 6; ++q) {
      const int i = PAIR_I[q], j = PAIR_J[q];
      real v;
      if (i == j)
        v = stencil_at(u, d, p, i, fw->w2[i], fw->nt);
      else
        v = mixed_at(u, d, p, i, j, fw->w1[i], fw->w1[j], fw->nt);
      Dp[(72 + 6 * s + q) * stride] = v;
    }
  }
}

void orc_deriv_bundle(const real* F, real* D, int Sx, int Sy, int Sz, const real* h3,
                      int order) {
  FDW fw;
  if (fd_weights(&fw, h3, order)) return;
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < Sx; ++x)
    for (int y = 0; y < Sy; ++y)
      for (int z = 0; z < Sz; ++z) {
        const int p[3] = {x, y, z};
        point_derivs(F, &d, &fw, p, D + lin(&d, x, y, z), d.P);
      }
}

static void unlin(const Dims* d, long idx, int p[3]) {
  p[2] = (int)(idx % d->S[2]);
  idx /= d->S[2];
  p[1] = (int)(idx % d->S[1]);
  p[0] = (int)(idx / d->S[1]);
}

void orc_deriv_points(const real* F, int Sx, int Sy, int Sz, const real* h3, int order,
                      long npts, const long* idx, real* Dp) {
  FDW fw;
  if (fd_weights(&fw, h3, order)) return;
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
#pragma omp parallel for schedule(static)
  for (long n = 0; n < npts; ++n) {
    int p[3];
    unlin(&d, idx[n], p);
    point_derivs(F, &d, &fw, p, Dp + n, npts);
  }
}

/* ---------------------------------------------------------------- algebra */

typedef struct {
  real eta, lam[4], lamf[2], cahd, dt, dxi, hssl, sigssl, t;
} Scal;

static void scal_from13(Scal* s, const real* sc) {
  s->eta = sc[0];
  for (int i = 0; i < 4; ++i) s->lam[i] = sc[1 + i];
  s->lamf[0] = sc[5];
  s->lamf[1] = sc[6];
  s->cahd = sc[7];
  s->dt = sc[8];
  s->dxi = sc[9];
  s->hssl = sc[10];
  s->sigssl = sc[11];
  s->t = sc[12];
}

/* S16 (eta, C, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz) -> 13 */
static void s16_to_13(const real* S, real* sc) {
  sc[0] = S[0];
  for (int i = 0; i < 4; ++i) sc[1 + i] = S[7 + i];
  sc[5] = S[11];
  sc[6] = S[12];
  sc[7] = S[1];
  sc[8] = S[2];
  sc[9] = S[3];
  sc[10] = S[4];
  sc[11] = S[5];
  sc[12] = S[6];
}

/* The BSSN right-hand side at one point from the 24 field values v[] and the 138
 * derivatives dv[] (layout as in the header). Output r[24] in state order.
 *
 * Equations (conformal variables chi, gt_ij, At_ij, K, Gt^i, lapse alpha, shift beta^i, B^i),
 * with d_k the centred FD derivative and gt^ij the inverse of gt_ij (adjugate / det):
 *   Gamma_kij  = (d_j gt_ki + d_i gt_kj - d_k gt_ij) / 2,   Gamma^m_ij = gt^mk Gamma_kij,
 *   CalGt^m    = gt^ij Gamma^m_ij,
 *   Gamma3^m_ij = Gamma^m_ij - (delta^m_i d_j chi + delta^m_j d_i chi - gt_ij gt^mk d_k chi)/(2chi)
 *   Rt_ij = -gt^lm d_l d_m gt_ij / 2 + (gt_ki d_j Gt^k + gt_kj d_i Gt^k) / 2
 *           + CalGt^k (Gamma_ijk + Gamma_jik) / 2
 *           + gt^lm (Gamma^k_li Gamma_jkm + Gamma^k_lj Gamma_ikm + Gamma^k_im Gamma_klj)
 *   Rchi_ij = (d_i d_j chi - Gamma^k_ji d_k chi)/(2chi) - d_i chi d_j chi/(4chi^2)
 *           + gt_ij [gt^kl (d_k d_l chi - 3 d_k chi d_l chi/(2chi)) - CalGt^m d_m chi]/(2chi)
 *   DDalpha_ij = d_i d_j alpha - Gamma3^l_ij d_l alpha
 *   H = chi gt^ij R_ij - At_ij At^ij + 2K^2/3
 *   d_t alpha = lambda0 beta.d alpha - 2 alpha K - sqrt(chi) h_ssl (alpha - sqrt(chi)) e^{-t^2/2s^2}
 *   d_t beta^i = 3/4 (lf0 + lf1 alpha) B^i + lambda1 beta.d beta^i
 *   d_t gt_ij = beta.d gt_ij + gt_ik d_j b^k + gt_kj d_i b^k - 2/3 gt_ij d_k b^k - 2 alpha At_ij
 *   d_t chi  = beta.d chi + 2/3 chi (alpha K - d_k b^k) + C dx^2 chi H / (dt (1 + 10 dx^2))
 *   d_t At_ij = beta.d At_ij + At_ik d_j b^k + At_kj d_i b^k - 2/3 At_ij d_k b^k
 *             + chi [alpha R_ij - DDalpha_ij]^TF + alpha (K At_ij - 2 At_ik gt^kl At_lj)
 *   d_t K = beta.d K - chi gt^ij DDalpha_ij + alpha (K^2/3 + At_ij At^ij)
 *   d_t Gt^i = beta.d Gt^i - CalGt^j d_j b^i + 2/3 CalGt^i d_j b^j
 *            + gt^jk d_j d_k b^i + gt^ij d_j d_k b^k / 3 - 2 At^ij d_j alpha
 *            + 2 alpha Gamma^i_jk At^jk - alpha (3/chi At^ij d_j chi + 4/3 gt^ij d_j K)
 *   d_t B^i = d_t Gt^i - eta B^i + lambda2 beta.d B^i - lambda3 beta.d Gt^i
 * (the CAHD+SSL variant the reference evaluates: vendor/bssneqs_SSL_HD_dxsq.cpp output lines
 * a_rhs .. B_rhs2; advection is centred, not upwinded).
 */
static void algebra_point(const real* v, const real* dv, const Scal* S, real* r,
                          real* ham_out, real* mom_out) {
#define D1(f, a) dv[3 * (f) + (a)]
#define D2(s, i, j) dv[72 + 6 * (s) + SYM[i][j]]
  /* second-derivative slots: 0 alpha, 1 chi, 2..7 gt0..5, 8..10 beta0..2 */
  const real al = v[ALPHA], chi = v[CHI], K = v[KK];
  real g[3][3], A[3][3], ig[3][3];
  real b[3], Gt[3], B[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      g[i][j] = v[GT0 + SYM[i][j]];
      A[i][j] = v[AT0 + SYM[i][j]];
    }
  for (int i = 0; i < 3; ++i) {
    b[i] = v[BETA0 + i];
    Gt[i] = v[GTI0 + i];
    B[i] = v[BB0 + i];
  }
  real db[3][3], dg[3][3][3], dchi[3], dal[3], dK[3], dGt[3][3];
  for (int j = 0; j < 3; ++j) {
    for (int i = 0; i < 3; ++i) db[j][i] = D1(BETA0 + i, j);      /* d_j beta^i */
    for (int k = 0; k < 3; ++k) dGt[j][k] = D1(GTI0 + k, j);      /* d_j Gt^k */
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) dg[j][i][k] = D1(GT0 + SYM[i][k], j); /* d_j gt_ik */
    dchi[j] = D1(CHI, j);
    dal[j] = D1(ALPHA, j);
    dK[j] = D1(KK, j);
  }
  /* inverse conformal metric */
  const real det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[1][2]) -
                     g[0][1] * (g[0][1] * g[2][2] - g[1][2] * g[0][2]) +
                     g[0][2] * (g[0][1] * g[1][2] - g[1][1] * g[0][2]);
  real adj[3][3];
  adj[0][0] = g[1][1] * g[2][2] - g[1][2] * g[1][2];
  adj[0][1] = g[0][2] * g[1][2] - g[0][1] * g[2][2];
  adj[0][2] = g[0][1] * g[1][2] - g[0][2] * g[1][1];
  adj[1][1] = g[0][0] * g[2][2] - g[0][2] * g[0][2];
  adj[1][2] = g[0][1] * g[0][2] - g[0][0] * g[1][2];
  adj[2][2] = g[0][0] * g[1][1] - g[0][1] * g[0][1];
  adj[1][0] = adj[0][1];
  adj[2][0] = adj[0][2];
  adj[2][1] = adj[1][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ig[i][j] = adj[i][j] / det;

  /* Christoffels */
  real C1[3][3][3], C2[3][3][3], C3[3][3][3], CalGt[3], igdchi[3];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        C1[k][i][j] = 0.5 * (dg[j][k][i] + dg[i][k][j] - dg[k][i][j]);
  for (int m = 0; m < 3; ++m)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        real s = 0.0;
        for (int k = 0; k < 3; ++k) s += ig[m][k] * C1[k][i][j];
        C2[m][i][j] = s;
      }
  for (int m = 0; m < 3; ++m) {
    real s = 0.0, t = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s += ig[i][j] * C2[m][i][j];
    CalGt[m] = s;
    for (int k = 0; k < 3; ++k) t += ig[m][k] * dchi[k];
    igdchi[m] = t;
  }
  for (int m = 0; m < 3; ++m)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        C3[m][i][j] = C2[m][i][j] - 0.5 / chi *
                                        ((m == i ? dchi[j] : 0.0) + (m == j ? dchi[i] : 0.0) -
                                         g[i][j] * igdchi[m]);

  /* conformal Ricci, metric part and chi part */
  real R[3][3];
  real lapchi = 0.0;
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      lapchi += ig[k][l] * (D2(1, k, l) - 1.5 / chi * dchi[k] * dchi[l]);
  for (int m = 0; m < 3; ++m) lapchi -= CalGt[m] * dchi[m];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real s = 0.0;
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) s += ig[l][m] * D2(2 + SYM[i][j], l, m);
      s *= -0.5;
      real t = 0.0;
      for (int k = 0; k < 3; ++k) t += g[k][i] * dGt[j][k] + g[k][j] * dGt[i][k];
      s += 0.5 * t;
      t = 0.0;
      for (int k = 0; k < 3; ++k) t += CalGt[k] * (C1[i][j][k] + C1[j][i][k]);
      s += 0.5 * t;
      t = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m)
            t += ig[l][m] * (C2[k][l][i] * C1[j][k][m] + C2[k][l][j] * C1[i][k][m] +
                             C2[k][i][m] * C1[k][l][j]);
      s += t;
      real c = 0.0;
      for (int k = 0; k < 3; ++k) c += C2[k][j][i] * dchi[k];
      const real rchi = 0.5 / chi * (D2(1, i, j) - c) - 0.25 / (chi * chi) * dchi[i] * dchi[j] +
                          0.5 / chi * g[i][j] * lapchi;
      R[i][j] = s + rchi;
    }

  real DDa[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real s = 0.0;
      for (int l = 0; l < 3; ++l) s += C3[l][i][j] * dal[l];
      DDa[i][j] = D2(0, i, j) - s;
    }
  const real divb = db[0][0] + db[1][1] + db[2][2];
  real Aup[3][3], AA = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real s = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) s += ig[i][k] * ig[j][l] * A[k][l];
      Aup[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) AA += A[i][j] * Aup[i][j];

#define ADV(f) (b[0] * D1(f, 0) + b[1] * D1(f, 1) + b[2] * D1(f, 2))
  /* lapse (with slice locking) */
  const real sq = SQRT(chi);
  r[ALPHA] = S->lam[0] * ADV(ALPHA) - 2.0 * al * K -
             sq * S->hssl * (al - sq) * EXP(-0.5 * S->t * S->t / (S->sigssl * S->sigssl));
  /* shift */
  for (int i = 0; i < 3; ++i)
    r[BETA0 + i] = 0.75 * (S->lamf[0] + S->lamf[1] * al) * B[i] + S->lam[1] * ADV(BETA0 + i);
  /* conformal metric */
  for (int q = 0; q < 6; ++q) {
    const int i = PAIR_I[q], j = PAIR_J[q];
    real s = 0.0;
    for (int k = 0; k < 3; ++k) s += g[i][k] * db[j][k] + g[k][j] * db[i][k];
    r[GT0 + q] = ADV(GT0 + q) + s - ((real)2 / 3) * g[i][j] * divb - 2.0 * al * A[i][j];
  }
  /* Hamiltonian constraint and the damped conformal factor */
  real Rs = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rs += ig[i][j] * R[i][j];
  const real ham = chi * Rs - AA + ((real)2 / 3) * K * K;
  const real dx2 = S->dxi * S->dxi;
  r[CHI] = ADV(CHI) + ((real)2 / 3) * chi * (al * K - divb) +
           S->cahd * dx2 * chi * ham / (S->dt * (10.0 * dx2 + 1.0));
  /* trace-free curvature */
  real X[3][3], trX = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) X[i][j] = al * R[i][j] - DDa[i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) trX += ig[i][j] * X[i][j];
  for (int q = 0; q < 6; ++q) {
    const int i = PAIR_I[q], j = PAIR_J[q];
    real lie = 0.0, aa = 0.0;
    for (int k = 0; k < 3; ++k) lie += A[i][k] * db[j][k] + A[k][j] * db[i][k];
    for (int k = 0; k < 3; ++k) {
      real t = 0.0;
      for (int l = 0; l < 3; ++l) t += ig[k][l] * A[l][j];
      aa += A[i][k] * t;
    }
    r[AT0 + q] = ADV(AT0 + q) + lie - ((real)2 / 3) * A[i][j] * divb +
                 chi * (X[i][j] - g[i][j] * trX / 3.0) + al * (K * A[i][j] - 2.0 * aa);
  }
  /* trace of the extrinsic curvature */
  real lap = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) lap += ig[i][j] * DDa[i][j];
  r[KK] = ADV(KK) - chi * lap + al * (K * K / 3.0 + AA);
  /* conformal connection functions and the Gamma-driver auxiliary */
  for (int i = 0; i < 3; ++i) {
    real s = 0.0;
    for (int j = 0; j < 3; ++j) s += b[j] * dGt[j][i];
    const real advGt = s;
    for (int j = 0; j < 3; ++j) s -= CalGt[j] * db[j][i];
    s += ((real)2 / 3) * CalGt[i] * divb;
    real t = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        t += ig[j][k] * D2(8 + i, j, k) + ig[i][j] * D2(8 + k, j, k) / 3.0;
    s += t;
    t = 0.0;
    for (int j = 0; j < 3; ++j) t += 2.0 * Aup[i][j] * dal[j];
    s -= t;
    t = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) t += 2.0 * al * C2[i][j][k] * Aup[j][k];
    s += t;
    t = 0.0;
    for (int j = 0; j < 3; ++j)
      t += al * (3.0 / chi * Aup[i][j] * dchi[j] + ((real)4 / 3) * ig[i][j] * dK[j]);
    s -= t;
    r[GTI0 + i] = s;
    r[BB0 + i] = s - S->eta * B[i] + S->lam[2] * ADV(BB0 + i) - S->lam[3] * advGt;
  }
  if (ham_out) *ham_out = ham;
  if (mom_out) {
    /* momentum constraint M^i = gt^jk (d_j At_ki - Gamma^m_ji At_km) - Gt^j At_ij
     *                           - 3/(2chi) gt^jk At_ki d_j chi - 2/3 d_i K */
    for (int i = 0; i < 3; ++i) {
      real s = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          real c = 0.0;
          for (int m = 0; m < 3; ++m) c += C2[m][j][i] * A[k][m];
          s += ig[j][k] * (D1(AT0 + SYM[k][i], j) - c);
        }
      real t = 0.0;
      for (int j = 0; j < 3; ++j) t += Gt[j] * A[i][j];
      s -= t;
      t = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) t += ig[j][k] * A[k][i] * dchi[j] / chi;
      s -= 1.5 * t;
      s -= ((real)2 / 3) * dK[i];
      mom_out[i] = s;
    }
  }
#undef ADV
#undef D1
#undef D2
}

void orc_algebra(long n, const real* F, const real* D, const real* sc13, real* OUT) {
  Scal S;
  scal_from13(&S, sc13);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; ++p) {
    real v[NF], dv[ND], r[NF];
    for (int f = 0; f < NF; ++f) v[f] = F[f * n + p];
    for (int d = 0; d < ND; ++d) dv[d] = D[d * n + p];
    algebra_point(v, dv, &S, r, NULL, NULL);
    for (int f = 0; f < NF; ++f) OUT[f * n + p] = r[f];
  }
}

/* Kreiss-Oliger sum over the three axes at one point for one field: ((0 + kx) + ky) + kz */
static real ko_at(const real* u, const Dims* d, const int p[3], real wko[3][9], int nt) {
  real acc = 0.0;
  for (int a = 0; a < 3; ++a) acc = acc + stencil_at(u, d, p, a, wko[a], nt);
  return acc;
}

/* The Kreiss-Oliger order paired with an FD order: order + 2 (an order-2m FD stencil with
 * order-2m+2 dissipation, common/src/mcs_common/derivatives.py:55-58), capped at 8, the widest
 * KO stencil the reference has (derivatives.py:45-49): FD order 8 pairs with KO order 8, the
 * reference's own default for it (derivatives.py:64). */
static int default_ko_order(int order) { return order + 2 < 8 ? order + 2 : 8; }

static int ko_weights(real wko[3][9], const real* h3, int ko_order, real sigma) {
  real c[9];
  const int nt = orc_stencil(2, ko_order, c);
  if (nt < 0) return -1;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < nt; ++k) wko[a][k] = c[k] * (sigma / h3[a]);
  return nt;
}

int orc_rhs_points(const real* F, int Sx, int Sy, int Sz, const real* S16, int order,
                   int ko_order, real ko_sigma, long npts, const long* idx, real* OUT) {
  FDW fw;
  const real h3[3] = {S16[13], S16[14], S16[15]};
  if (ko_order == 0) ko_order = default_ko_order(order);
  if (fd_weights(&fw, h3, order)) return -1;
  real wko[3][9];
  const int nko = ko_weights(wko, h3, ko_order, ko_sigma);
  if (nko < 0) return -1;
  real sc[13];
  s16_to_13(S16, sc);
  Scal S;
  scal_from13(&S, sc);
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
#pragma omp parallel for schedule(static)
  for (long n = 0; n < npts; ++n) {
    int p[3];
    unlin(&d, idx[n], p);
    real v[NF], dv[ND], r[NF];
    for (int f = 0; f < NF; ++f) v[f] = F[f * d.P + idx[n]];
    point_derivs(F, &d, &fw, p, dv, 1);
    algebra_point(v, dv, &S, r, NULL, NULL);
    if (ko_sigma > 0.0)
      for (int f = 0; f < NF; ++f) r[f] = r[f] + ko_at(F + f * d.P, &d, p, wko, nko);
    for (int f = 0; f < NF; ++f) OUT[f * npts + n] = r[f];
  }
  return 0;
}

int orc_rhs(const real* F, real* OUT, int Sx, int Sy, int Sz, const real* S16, int order,
            int ko_order, real ko_sigma) {
  FDW fw;
  const real h3[3] = {S16[13], S16[14], S16[15]};
  if (ko_order == 0) ko_order = default_ko_order(order);
  if (fd_weights(&fw, h3, order)) return -1;
  real wko[3][9];
  const int nko = ko_weights(wko, h3, ko_order, ko_sigma);
  if (nko < 0) return -1;
  real sc[13];
  s16_to_13(S16, sc);
  Scal S;
  scal_from13(&S, sc);
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < Sx; ++x)
    for (int y = 0; y < Sy; ++y)
      for (int z = 0; z < Sz; ++z) {
        const int p[3] = {x, y, z};
        const long q = lin(&d, x, y, z);
        real v[NF], dv[ND], r[NF];
        for (int f = 0; f < NF; ++f) v[f] = F[f * d.P + q];
        point_derivs(F, &d, &fw, p, dv, 1);
        algebra_point(v, dv, &S, r, NULL, NULL);
        if (ko_sigma > 0.0)
          for (int f = 0; f < NF; ++f) r[f] = r[f] + ko_at(F + f * d.P, &d, p, wko, nko);
        for (int f = 0; f < NF; ++f) OUT[f * d.P + q] = r[f];
      }
  return 0;
}

/* ---------------------------------------------------------------- constraints */

void orc_constraints_points(long n, const real* F, const real* D, real* HM) {
  Scal S;
  memset(&S, 0, sizeof S);
  S.dt = 1.0;
  S.sigssl = 1.0;
#pragma omp parallel for schedule(static)
  for (long p = 0; p < n; ++p) {
    real v[NF], dv[ND], r[NF], ham, mom[3];
    for (int f = 0; f < NF; ++f) v[f] = F[f * n + p];
    for (int d = 0; d < ND; ++d) dv[d] = D[d * n + p];
    algebra_point(v, dv, &S, r, &ham, mom);
    HM[p] = ham;
    for (int i = 0; i < 3; ++i) HM[(1 + i) * n + p] = mom[i];
  }
}

void orc_constraints(const real* F, real* HM, int Sx, int Sy, int Sz, const real* h3,
                     int order) {
  FDW fw;
  if (fd_weights(&fw, h3, order)) return;
  Dims d = {{Sx, Sy, Sz}, (long)Sx * Sy * Sz};
  Scal S;
  memset(&S, 0, sizeof S);
  S.dt = 1.0;
  S.sigssl = 1.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int x = 0; x < Sx; ++x)
    for (int y = 0; y < Sy; ++y)
      for (int z = 0; z < Sz; ++z) {
        const int p[3] = {x, y, z};
        const long q = lin(&d, x, y, z);
        real v[NF], dv[ND], r[NF], ham, mom[3];
        for (int f = 0; f < NF; ++f) v[f] = F[f * d.P + q];
        point_derivs(F, &d, &fw, p, dv, 1);
        algebra_point(v, dv, &S, r, &ham, mom);
        HM[q] = ham;
        for (int i = 0; i < 3; ++i) HM[(1 + i) * d.P + q] = mom[i];
      }
}

void orc_constraint_l2(const real* F, int Sx, int Sy, int Sz, const real* h3, int order,
                       int ng, real* out2) {
  const long P = (long)Sx * Sy * Sz;
  real* HM = (real*)malloc(sizeof(real) * 4 * P);
  if (!HM) {
    out2[0] = out2[1] = NAN;
    return;
  }
  orc_constraints(F, HM, Sx, Sy, Sz, h3, order);
  real sh = 0.0, sm = 0.0;
  long cnt = 0;
  for (int x = ng; x < Sx - ng; ++x)
    for (int y = ng; y < Sy - ng; ++y)
      for (int z = ng; z < Sz - ng; ++z) {
        const long q = ((long)x * Sy + y) * Sz + z;
        sh += HM[q] * HM[q];
        sm += HM[P + q] * HM[P + q] + HM[2 * P + q] * HM[2 * P + q] + HM[3 * P + q] * HM[3 * P + q];
        ++cnt;
      }
  out2[0] = SQRT(sh / (real)cnt);
  out2[1] = SQRT(sm / (real)cnt);
  free(HM);
}

/* ---------------------------------------------------------------- step */

void orc_periodic_sync(real* F, int Sx, int Sy, int Sz, int ng) {
  const long P = (long)Sx * Sy * Sz;
  const int S[3] = {Sx, Sy, Sz};
  /* x, then y, then z; each copies ng planes from the opposite interior
   * (data[:ng] = data[-2ng:-ng]; data[-ng:] = data[ng:2ng]) */
  for (int axis = 0; axis < 3; ++axis) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < NF; ++f) {
      real* u = F + f * P;
      for (int x = 0; x < Sx; ++x)
        for (int y = 0; y < Sy; ++y)
          for (int z = 0; z < Sz; ++z) {
            int p[3] = {x, y, z};
            const int c = p[axis];
            int src;
            if (c < ng)
              src = c + S[axis] - 2 * ng;
            else if (c >= S[axis] - ng)
              src = c - S[axis] + 2 * ng;
            else
              continue;
            p[axis] = src;
            u[((long)x * Sy + y) * Sz + z] = u[((long)p[0] * Sy + p[1]) * Sz + p[2]];
          }
    }
  }
}

void orc_enforce(real* F, long P, real chi_floor, real alpha_floor) {
#pragma omp parallel for schedule(static)
  for (long q = 0; q < P; ++q) {
    real g[6], a[6];
    for (int i = 0; i < 6; ++i) g[i] = F[(GT0 + i) * P + q];
    const real det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
                       g[2] * (g[1] * g[4] - g[3] * g[2]);
    const real scale = POW(det, -((real)1 / 3));
    for (int i = 0; i < 6; ++i) g[i] = g[i] * scale;
    const real iuxx = g[3] * g[5] - g[4] * g[4];
    const real iuxy = g[2] * g[4] - g[1] * g[5];
    const real iuxz = g[1] * g[4] - g[2] * g[3];
    const real iuyy = g[0] * g[5] - g[2] * g[2];
    const real iuyz = g[1] * g[2] - g[0] * g[4];
    const real iuzz = g[0] * g[3] - g[1] * g[1];
    for (int i = 0; i < 6; ++i) a[i] = F[(AT0 + i) * P + q];
    const real trA = iuxx * a[0] + iuyy * a[3] + iuzz * a[5] +
                       2.0 * (iuxy * a[1] + iuxz * a[2] + iuyz * a[4]);
    const real third = trA / 3.0;
    for (int i = 0; i < 6; ++i) {
      F[(GT0 + i) * P + q] = g[i];
      F[(AT0 + i) * P + q] = a[i] - g[i] * third;
    }
    if (F[CHI * P + q] < chi_floor) F[CHI * P + q] = chi_floor;
    if (F[ALPHA * P + q] < alpha_floor) F[ALPHA * P + q] = alpha_floor;
  }
}

int orc_stage_rhs(const real* Y, real* Kout, real* scratch, int Sx, int Sy, int Sz,
                  const real* S16, int order, int ko_order, real ko_sigma, int ng,
                  int periodic, real t) {
  const long P = (long)Sx * Sy * Sz;
  real S[16];
  memcpy(S, S16, sizeof S);
  S[6] = t;
  const real* src = Y;
  if (periodic) {
    memcpy(scratch, Y, sizeof(real) * NF * P);
    orc_periodic_sync(scratch, Sx, Sy, Sz, ng);
    src = scratch;
  }
  return orc_rhs(src, Kout, Sx, Sy, Sz, S, order, ko_order, ko_sigma);
}

int orc_rk4_step(real* U, int Sx, int Sy, int Sz, const real* S16, int order, int ko_order,
                 real ko_sigma, int ng, int periodic, real t, real dt, real chi_floor,
                 real alpha_floor, real* work) {
  const long N = (long)NF * Sx * Sy * Sz;
  real *k1 = work, *k2 = work + N, *k3 = work + 2 * N, *k4 = work + 3 * N;
  real *Y = work + 4 * N, *scr = work + 5 * N;
  real S[16];
  memcpy(S, S16, sizeof S);
  S[2] = dt; /* the CAHD factor uses the full step at every stage (evolve.py:121-124) */
  const real hdt = 0.5 * dt;
  int rc = orc_stage_rhs(U, k1, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t);
  if (rc) return rc;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) Y[i] = U[i] + hdt * k1[i];
  orc_stage_rhs(Y, k2, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t + hdt);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) Y[i] = U[i] + hdt * k2[i];
  orc_stage_rhs(Y, k3, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t + hdt);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) Y[i] = U[i] + dt * k3[i];
  orc_stage_rhs(Y, k4, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t + dt);
  const real w = dt / 6.0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) U[i] = U[i] + w * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
  orc_enforce(U, (long)Sx * Sy * Sz, chi_floor, alpha_floor);
  return 0;
}

/* rk4_2_2 coefficients: Sanches, Brandt, Kalinani, Ji & Schnetter (arXiv:2603.05763), Table 1,
 * solution (2), as the reference enters them (2D/src/mcs2d/msrk.py:47-52). */
static const real MS_A20 = 1309.0 / 15500.0, MS_A21 = -31999.0 / 15500.0;
static const real MS_A30 = -241289.0 / 5880000.0, MS_A31 = 22846301.0 / 16170000.0,
                    MS_A32 = -936169.0 / 2587200.0;
static const real MS_B0 = -191.0 / 882.0, MS_B1 = 48241.0 / 59994.0,
                    MS_B2 = 193750.0 / 4351347.0, MS_B3 = 100000.0 / 271791.0;

void orc_msrk_coeffs(real* c) {
  c[0] = MS_A20; c[1] = MS_A21; c[2] = MS_A30; c[3] = MS_A31; c[4] = MS_A32;
  c[5] = MS_B0; c[6] = MS_B1; c[7] = MS_B2; c[8] = MS_B3;
}

int orc_msrk_step(real* U, real* fprev, int Sx, int Sy, int Sz, const real* S16,
                  int order, int ko_order, real ko_sigma, int ng, int periodic, real t,
                  real dt, real chi_floor, real alpha_floor, real* work) {
  const long N = (long)NF * Sx * Sy * Sz;
  real *k1 = work, *k2 = work + N, *k3 = work + 2 * N, *Y = work + 3 * N, *scr = work + 4 * N;
  real S[16];
  memcpy(S, S16, sizeof S);
  S[2] = dt;
  const real c2 = MS_A20 + MS_A21;
  const real c3 = MS_A30 + MS_A31 + MS_A32;
  const real* k0 = fprev;
  int rc = orc_stage_rhs(U, k1, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t);
  if (rc) return rc;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) Y[i] = U[i] + dt * (MS_A20 * k0[i] + MS_A21 * k1[i]);
  orc_stage_rhs(Y, k2, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t + c2 * dt);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i)
    Y[i] = U[i] + dt * (MS_A30 * k0[i] + MS_A31 * k1[i] + MS_A32 * k2[i]);
  orc_stage_rhs(Y, k3, scr, Sx, Sy, Sz, S, order, ko_order, ko_sigma, ng, periodic, t + c3 * dt);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) {
    U[i] = U[i] + dt * (MS_B0 * k0[i] + MS_B1 * k1[i] + MS_B2 * k2[i] + MS_B3 * k3[i]);
    fprev[i] = k1[i];
  }
  orc_enforce(U, (long)Sx * Sy * Sz, chi_floor, alpha_floor);
  return 0;
}
