/*
 * CPU oracle for the BSSN right-hand side and its RK4 / MSRK time step.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's verbatim path,
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (integer-schemes_amd/) links, loads or calls it.
 *
 * Array conventions (shared with the reference and the HIP library):
 *   state  F[f*P + (x*Sy + y)*Sz + z], P = Sx*Sy*Sz, 24 fields in the order
 *          alpha, chi, K, gt0..5, beta0..2, At0..5, Gt0..2, B0..2
 *          (3D/src/bssn3d/state.py:26-44).
 *   bundle D[d*P + p], 138 derivatives: d = 3*f + axis (72 first derivatives, all fields),
 *          d = 72 + 6*s + pair (66 second derivatives) with s over
 *          (alpha, chi, gt0..gt5, beta0..beta2) and pair over (00,01,02,11,12,22).
 *   S16    the reference's scalar pack (3D/src/bssn3d/fused_rhs_cuda.py:52-55):
 *          eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz.
 */
#ifndef BSSN_ORACLE_H
#define BSSN_ORACLE_H

/* Working precision: double by default; the long-double build (liboracle_ld.so, the
 * accuracy reference of the tests) compiles the same source with ORC_REAL=long double. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#ifdef __cplusplus
extern "C" {
#endif

int orc_version(void);
void orc_set_threads(int n);
int orc_get_threads(void);

/* Stencil coefficients (3D common/src/mcs_common/derivatives.py:33-49). kind 0 = first
 * derivative, 1 = second derivative, 2 = Kreiss-Oliger. Writes order+1 values, returns the
 * count, or -1 for an unsupported order. */
int orc_stencil(int kind, int order, real* coeffs);

/* One-axis operator over a whole (Sx,Sy,Sz) array with edge padding by the stencil's own
 * reach (derivatives.py:78-88). factor = 1/h (d1), 1/h^2 (d2) or sigma/h (KO). */
void orc_apply_axis(const real* u, real* out, int Sx, int Sy, int Sz, int axis,
                    int kind, int order, real factor);

/* The 138-derivative bundle at every padded point (derivative_bundle.py:31-48). */
void orc_deriv_bundle(const real* F, real* D, int Sx, int Sy, int Sz,
                      const real* h3, int order);

/* The 138 derivatives at a list of points only (same arithmetic as the bundle). */
void orc_deriv_points(const real* F, int Sx, int Sy, int Sz, const real* h3, int order,
                      long npts, const long* idx, real* Dp /* 138*npts, d-major */);

/* Pointwise BSSN algebra over n points. sc13 = eta, lambda0..3, lambda_f0..1, C_CAHD, dt,
 * dx_i, h_ssl, sig_ssl, t. */
void orc_algebra(long n, const real* F, const real* D, const real* sc13, real* OUT);

/* Full padded RHS: bundle + algebra (rhs.py:102-139), optional KO added when ko_sigma > 0
 * (evolve.py:76-90). ko_order 0 selects the default pairing (order + 2, at most 8). */
int orc_rhs(const real* F, real* OUT, int Sx, int Sy, int Sz, const real* S16,
            int order, int ko_order, real ko_sigma);

/* RHS (+KO) at listed points only. OUT is 24*npts, field-major. */
int orc_rhs_points(const real* F, int Sx, int Sy, int Sz, const real* S16, int order,
                   int ko_order, real ko_sigma, long npts, const long* idx, real* OUT);

/* Hamiltonian + momentum constraints at every padded point (constraints.py:38-51):
 * HM = 4 arrays (H, M0, M1, M2). */
void orc_constraints(const real* F, real* HM, int Sx, int Sy, int Sz, const real* h3,
                     int order);

/* Constraints from given fields and derivatives at n points: HM is 4*n. */
void orc_constraints_points(long n, const real* F, const real* D, real* HM);

/* Interior RMS norms (constraints.py:53-62): out2 = {H_l2, M_l2}. */
void orc_constraint_l2(const real* F, int Sx, int Sy, int Sz, const real* h3, int order,
                       int ng, real* out2);

void orc_periodic_sync(real* F, int Sx, int Sy, int Sz, int ng);
void orc_enforce(real* F, long P, real chi_floor, real alpha_floor);

/* One RK4 step in place (evolve.py:125-132), with periodic ghost sync before every RHS
 * evaluation when periodic != 0. work must hold 6 states. */
int orc_rk4_step(real* U, int Sx, int Sy, int Sz, const real* S16, int order, int ko_order,
                 real ko_sigma, int ng, int periodic, real t, real dt,
                 real chi_floor, real alpha_floor, real* work);

/* The rk4_2_2 coefficients the step uses: a20, a21, a30, a31, a32, b0, b1, b2, b3. */
void orc_msrk_coeffs(real* c);

/* One MSRK rk4_2_2 step (evolve.py:139-153). fprev holds f(y_{n-1}) on entry and f(y_n) on
 * exit. work must hold 6 states. */
int orc_msrk_step(real* U, real* fprev, int Sx, int Sy, int Sz, const real* S16,
                  int order, int ko_order, real ko_sigma, int ng, int periodic, real t,
                  real dt, real chi_floor, real alpha_floor, real* work);

/* Full RHS (+KO, +sync) used as the stage function f(y, t) by the steppers. */
int orc_stage_rhs(const real* Y, real* K, real* scratch, int Sx, int Sy, int Sz,
                  const real* S16, int order, int ko_order, real ko_sigma, int ng,
                  int periodic, real t);

#ifdef __cplusplus
}
#endif
#endif
