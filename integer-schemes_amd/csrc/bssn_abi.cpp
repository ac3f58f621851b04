// C ABI of the BSSN HIP library: argument checks, scalar unpacking and kernel sequencing.
// Declarations and their reference counterparts: include/bssn_hip.h.
#include "bssn_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "bssn_kernels.h"
#include "bssn_params.h"

using namespace bssn;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return BSSN_OK;
  return fail(BSSN_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

bool valid_order(int ord) { return ord == 4 || ord == 6; }

// Scales of the integer stencils (bssn_brick.h): the denominators of the centred and
// Kreiss-Oliger coefficients (common/src/mcs_common/derivatives.py:33-49) times the spacings.
Scales make_scales(const double h3[3], int ord, double sigma) {
  Scales s;
  const double den1 = ord == 4 ? 12.0 : 60.0;    // first derivative
  const double den2 = ord == 4 ? 12.0 : 180.0;   // second derivative
  const double denk = ord == 4 ? 64.0 : 256.0;   // KO of order ord + 2
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

// S16 = eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz
Scal make_scal(const double* S, double dt, double t) {
  Scal s;
  s.eta = S[0];
  for (int i = 0; i < 4; ++i) s.lam[i] = S[7 + i];
  s.lamf[0] = S[11];
  s.lamf[1] = S[12];
  s.ssl_h = S[4];
  s.ssl_e = std::exp(-0.5 * t * t / (S[5] * S[5]));
  const double dx2 = S[3] * S[3];
  s.cahd = S[1] * dx2 / (dt * (10.0 * dx2 + 1.0));
  return s;
}

Dims full_dims(int Sx, int Sy, int Sz, int ng) {
  Dims d;
  d.S[0] = Sx; d.S[1] = Sy; d.S[2] = Sz;
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = 0;
    d.n[a] = d.S[a];
  }
  d.ng = ng;
  d.P = (long)Sx * Sy * Sz;
  return d;
}

Dims interior_dims(int Sx, int Sy, int Sz, int ng) {
  Dims d = full_dims(Sx, Sy, Sz, ng);
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = ng;
    d.n[a] = d.S[a] - 2 * ng;
  }
  return d;
}

int check_shape(int Sx, int Sy, int Sz) {
  if (Sx < 1 || Sy < 1 || Sz < 1)
    return fail(BSSN_E_ARG, "grid extents must be positive");
  const double P = (double)Sx * Sy * Sz;
  if (P * 138.0 > 9.2e18 || P > 2.0e9)
    return fail(BSSN_E_ARG, "grid too large for 32-bit point offsets");
  return BSSN_OK;
}

int check_h(const double* h3) {
  if (!h3) return fail(BSSN_E_ARG, "null spacing array");
  for (int a = 0; a < 3; ++a)
    if (!(h3[a] > 0.0) || !std::isfinite(h3[a])) return fail(BSSN_E_ARG, "spacings must be > 0");
  return BSSN_OK;
}

int check_s16(const double* S) {
  if (!S) return fail(BSSN_E_ARG, "null scalar pack S16");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(S[i])) return fail(BSSN_E_ARG, "non-finite entry in S16");
  if (!(S[13] > 0.0 && S[14] > 0.0 && S[15] > 0.0))
    return fail(BSSN_E_ARG, "S16 spacings (entries 13..15) must be > 0");
  if (S[2] == 0.0) return fail(BSSN_E_ARG, "S16 dt (entry 2) must be nonzero (CAHD divides by it)");
  if (S[5] == 0.0) return fail(BSSN_E_ARG, "S16 sig_ssl (entry 5) must be nonzero");
  return BSSN_OK;
}

int check_ng(int Sx, int Sy, int Sz, int ng) {
  if (ng < 0) return fail(BSSN_E_ARG, "ng must be >= 0");
  if (Sx - 2 * ng < ng || Sy - 2 * ng < ng || Sz - 2 * ng < ng || Sx - 2 * ng < 1 ||
      Sy - 2 * ng < 1 || Sz - 2 * ng < 1)
    return fail(BSSN_E_ARG, "interior must hold at least max(1, ng) points per axis");
  return BSSN_OK;
}

#define TRY(x)                   \
  do {                           \
    const int rc_ = (x);         \
    if (rc_ != BSSN_OK) return rc_; \
  } while (0)

int run_stage(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc, int ord,
              bool ko, int mode, hipStream_t st) {
  return check_hip(launch_fused(a, d, w, sc, ord, ko, mode, st), "fused stage launch");
}

}  // namespace

extern "C" {

int bssn_abi_version(void) { return BSSN_HIP_ABI_VERSION; }

const char* bssn_last_error(void) { return g_err.c_str(); }

int bssn_rhs_f64(const double* F, const double* S16, double* OUT, int Sx, int Sy, int Sz,
                 int fd_order, double ko_sigma, void* stream) {
  if (!F || !OUT) return fail(BSSN_E_ARG, "null state pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_s16(S16));
  const double h3[3] = {S16[13], S16[14], S16[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const Scal sc = make_scal(S16, S16[2], S16[6]);
  const Dims d = full_dims(Sx, Sy, Sz, 0);
  StageArgs a{};
  a.Yin = F;
  a.Y = OUT;
  a.kind = KIND_RHS;
  return run_stage(a, d, w, sc, fd_order, ko_sigma > 0.0, CLAMP, (hipStream_t)stream);
}

int bssn_stage_rhs_f64(const double* Y, double* OUT, int Sx, int Sy, int Sz, const double* S16,
                       int fd_order, double ko_sigma, int ng, int periodic, void* stream) {
  if (!Y || !OUT) return fail(BSSN_E_ARG, "null state pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_s16(S16));
  if (periodic) TRY(check_ng(Sx, Sy, Sz, ng));
  const double h3[3] = {S16[13], S16[14], S16[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const Scal sc = make_scal(S16, S16[2], S16[6]);
  const Dims d = periodic ? interior_dims(Sx, Sy, Sz, ng) : full_dims(Sx, Sy, Sz, ng);
  StageArgs a{};
  a.Yin = Y;
  a.Y = OUT;
  a.kind = KIND_RHS;
  return run_stage(a, d, w, sc, fd_order, ko_sigma > 0.0, periodic ? WRAP : CLAMP,
                   (hipStream_t)stream);
}

int bssn_deriv_bundle_f64(const double* F, double* D, int Sx, int Sy, int Sz, int fd_order,
                          const double* h3, void* stream) {
  if (!F || !D) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  const Scales w = make_scales(h3, fd_order, 0.0);
  return check_hip(launch_deriv_bundle(F, D, full_dims(Sx, Sy, Sz, 0), w, fd_order,
                                       (hipStream_t)stream),
                   "deriv bundle launch");
}

int bssn_algebra_f64(const double* F, const double* D, const double* S16, double* OUT, long npts,
                     void* stream) {
  if (!F || !D || !OUT) return fail(BSSN_E_ARG, "null pointer");
  if (npts < 1) return fail(BSSN_E_ARG, "npts must be >= 1");
  TRY(check_s16(S16));
  const Scal sc = make_scal(S16, S16[2], S16[6]);
  return check_hip(launch_algebra(F, D, OUT, npts, sc, (hipStream_t)stream), "algebra launch");
}

int bssn_ko_add_f64(const double* F, double* OUT, int Sx, int Sy, int Sz, int fd_order,
                    double sigma, const double* h3, int periodic, int ng, void* stream) {
  if (!F || !OUT) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  if (periodic) TRY(check_ng(Sx, Sy, Sz, ng));
  if (!(sigma > 0.0)) return BSSN_OK;   // the reference adds 0 for sigma <= 0
  const Scales w = make_scales(h3, fd_order, sigma);
  const Dims d = periodic ? interior_dims(Sx, Sy, Sz, ng) : full_dims(Sx, Sy, Sz, ng);
  return check_hip(launch_ko_add(F, OUT, d, w, fd_order, periodic ? WRAP : CLAMP,
                                 (hipStream_t)stream),
                   "KO launch");
}

int bssn_periodic_sync_f64(double* F, int Sx, int Sy, int Sz, int ng, void* stream) {
  if (!F) return fail(BSSN_E_ARG, "null pointer");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_ng(Sx, Sy, Sz, ng));
  if (ng == 0) return BSSN_OK;
  return check_hip(launch_periodic_sync(F, full_dims(Sx, Sy, Sz, ng), (hipStream_t)stream),
                   "periodic sync launch");
}

int bssn_enforce_f64(double* F, long npts, double chi_floor, double alpha_floor, void* stream) {
  if (!F) return fail(BSSN_E_ARG, "null pointer");
  if (npts < 1) return fail(BSSN_E_ARG, "npts must be >= 1");
  return check_hip(launch_enforce(F, npts, chi_floor, alpha_floor, (hipStream_t)stream),
                   "enforce launch");
}

size_t bssn_step_workspace_bytes(int integrator, int Sx, int Sy, int Sz) {
  if (Sx < 1 || Sy < 1 || Sz < 1) return 0;
  const size_t state = (size_t)BSSN_NF * Sx * Sy * Sz * sizeof(double);
  return integrator == 0 ? 3 * state : 2 * state;
}

int bssn_rk4_step_f64(double* U, double* work, int Sx, int Sy, int Sz, const double* S16,
                      int fd_order, double ko_sigma, int ng, int periodic, double t, double dt,
                      double chi_floor, double alpha_floor, void* stream) {
  if (!U || !work) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(t))
    return fail(BSSN_E_ARG, "dt must be finite and nonzero, t finite");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_s16(S16));
  if (periodic) TRY(check_ng(Sx, Sy, Sz, ng));
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)Sx * Sy * Sz;
  const size_t NS = (size_t)BSSN_NF * P;
  double* Ya = work;
  double* Yb = work + NS;
  double* KS = work + 2 * NS;
  const double h3[3] = {S16[13], S16[14], S16[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const Dims d = periodic ? interior_dims(Sx, Sy, Sz, ng) : full_dims(Sx, Sy, Sz, ng);
  const int mode = periodic ? WRAP : CLAMP;
  const bool ko = ko_sigma > 0.0;
  const double hdt = 0.5 * dt;
  StageArgs a{};
  a.U = U;
  a.K1 = KS;
  a.Uout = U;
  a.dt = dt;
  a.chi_floor = chi_floor;
  a.alpha_floor = alpha_floor;
  // k1 = f(y, t);          Y = y + (dt/2) k1
  a.Yin = U; a.Y = Ya; a.kind = KIND_RK1; a.c0 = hdt;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t), fd_order, ko, mode, st));
  // k2 = f(Y, t + dt/2);   Y = y + (dt/2) k2
  a.Yin = Ya; a.Y = Yb; a.kind = KIND_RK2; a.c0 = hdt;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t + hdt), fd_order, ko, mode, st));
  // k3 = f(Y, t + dt/2);   Y = y + dt k3
  a.Yin = Yb; a.Y = Ya; a.kind = KIND_RK3; a.c0 = dt;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t + hdt), fd_order, ko, mode, st));
  // k4 = f(Y, t + dt);     y + (dt/6)(k1 + 2k2 + 2k3 + k4), then enforce
  a.Yin = Ya; a.Y = nullptr; a.kind = KIND_RK4; a.c0 = dt / 6.0;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t + dt), fd_order, ko, mode, st));
  if (periodic && ng > 0)
    return check_hip(launch_periodic_sync(U, full_dims(Sx, Sy, Sz, ng), st), "periodic sync");
  return BSSN_OK;
}

int bssn_msrk_step_f64(double* U, double* fprev, double* fnext, double* work, int Sx, int Sy,
                       int Sz, const double* S16, int fd_order, double ko_sigma, int ng,
                       int periodic, double t, double dt, double chi_floor, double alpha_floor,
                       void* stream) {
  if (!U || !fprev || !fnext || !work) return fail(BSSN_E_ARG, "null pointer");
  if (fprev == fnext) return fail(BSSN_E_ARG, "fprev and fnext must be distinct buffers");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(t))
    return fail(BSSN_E_ARG, "dt must be finite and nonzero, t finite");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_s16(S16));
  if (periodic) TRY(check_ng(Sx, Sy, Sz, ng));
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)Sx * Sy * Sz;
  const size_t NS = (size_t)BSSN_NF * P;
  double* Ya = work;
  double* Yb = work + NS;
  const double h3[3] = {S16[13], S16[14], S16[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const Dims d = periodic ? interior_dims(Sx, Sy, Sz, ng) : full_dims(Sx, Sy, Sz, ng);
  const int mode = periodic ? WRAP : CLAMP;
  const bool ko = ko_sigma > 0.0;
  const double c2 = MS_A20 + MS_A21;
  const double c3 = MS_A30 + MS_A31 + MS_A32;
  StageArgs a{};
  a.U = U;
  a.FP = fprev;
  a.K1 = fnext;
  a.Uout = U;
  a.dt = dt;
  a.chi_floor = chi_floor;
  a.alpha_floor = alpha_floor;
  // k1 = f(y, t);            Y2 = y + dt (a20 k0 + a21 k1)
  a.Yin = U; a.Y = Ya; a.kind = KIND_MS1; a.c0 = MS_A20; a.c1 = MS_A21;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t), fd_order, ko, mode, st));
  // k2 = f(Y2, t + c2 dt);   Y3 = y + dt (a30 k0 + a31 k1 + a32 k2)
  a.Yin = Ya; a.Y = Yb; a.kind = KIND_MS2; a.c0 = MS_A30; a.c1 = MS_A31; a.c2 = MS_A32;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t + c2 * dt), fd_order, ko, mode, st));
  // k3 = f(Y3, t + c3 dt);   y + dt (b0 k0 + b1 k1 + b2 k2 + b3 k3), then enforce
  a.Yin = Yb; a.Y = nullptr; a.kind = KIND_MS3;
  TRY(run_stage(a, d, w, make_scal(S16, dt, t + c3 * dt), fd_order, ko, mode, st));
  if (periodic && ng > 0)
    return check_hip(launch_periodic_sync(U, full_dims(Sx, Sy, Sz, ng), st), "periodic sync");
  return BSSN_OK;
}

int bssn_constraints_f64(const double* F, double* HM, int Sx, int Sy, int Sz, int fd_order,
                         const double* h3, void* stream) {
  if (!F || !HM) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4 or 6");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  const Scales w = make_scales(h3, fd_order, 0.0);
  return check_hip(launch_constraints(F, HM, full_dims(Sx, Sy, Sz, 0), w, fd_order,
                                      (hipStream_t)stream),
                   "constraints launch");
}

long bssn_l2_partial_doubles(int Sx, int Sy, int Sz, int ng) {
  (void)Sz;
  const long nx = Sx - 2 * ng, ny = Sy - 2 * ng;
  return nx > 0 && ny > 0 ? 2 * nx * ny : 0;
}

int bssn_constraint_l2_f64(const double* F, double* HM, double* partial, int Sx, int Sy, int Sz,
                           int fd_order, const double* h3, int ng, double* out2, void* stream) {
  if (!partial || !out2) return fail(BSSN_E_ARG, "null pointer");
  TRY(check_ng(Sx, Sy, Sz, ng));
  TRY(bssn_constraints_f64(F, HM, Sx, Sy, Sz, fd_order, h3, stream));
  hipStream_t st = (hipStream_t)stream;
  const Dims d = interior_dims(Sx, Sy, Sz, ng);
  const int nblk = sumsq_blocks(d);
  TRY(check_hip(launch_sumsq(HM, d, partial, nblk, st), "sumsq launch"));
  std::vector<double> host((size_t)2 * nblk);
  TRY(check_hip(hipMemcpyAsync(host.data(), partial, sizeof(double) * 2 * nblk,
                               hipMemcpyDeviceToHost, st),
                "copy partial sums"));
  TRY(check_hip(hipStreamSynchronize(st), "stream synchronize"));
  double sh = 0.0, sm = 0.0;
  for (int b = 0; b < nblk; ++b) {
    sh += host[2 * b];
    sm += host[2 * b + 1];
  }
  const double cnt = (double)d.n[0] * d.n[1] * d.n[2];
  out2[0] = std::sqrt(sh / cnt);
  out2[1] = std::sqrt(sm / cnt);
  return BSSN_OK;
}

}  // extern "C"
