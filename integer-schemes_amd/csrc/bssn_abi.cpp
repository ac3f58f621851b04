// C ABI of the BSSN HIP library: argument checks, scalar unpacking and kernel sequencing.
// Declarations and their reference counterparts: include/bssn_hip.h.
#include "bssn_hip.h"
#include "bssn_entries.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
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

bool valid_order(int ord) { return ord == 4 || ord == 6 || ord == 8; }
// Kreiss-Oliger order paired with an FD order: ord + 2, at most 8 (the reference's widest)
int ko_order_of(int ord) { return ord + 2 < 8 ? ord + 2 : 8; }

// S16 = eta, C_CAHD, dt, dx_i, h_ssl, sig_ssl, t, lambda0..3, lambda_f0..1, dx, dy, dz
Scal make_scal(const double* S, double dt, double t) {
  Scal s;
  s.eta = S[0];
  for (int i = 0; i < 4; ++i) s.lam[i] = S[7 + i];
  s.lamf[0] = S[11];
  s.lamf[1] = S[12];
  s.ssl_h = S[4];
  s.ssl_t = t;
  s.ssl_sig = S[5];
  s.ssl_e = 0.0;   // formed on the device from ssl_t and ssl_sig
  s.cahd = cahd_factor(S[1], S[3], dt);
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

// Ghost handling of a step (the "periodic" argument of the step entry points):
//   0 = edge: every padded point advanced with edge-padded stencils, no sync;
//   1 = wrapped: interior advanced with periodically wrapped stencils, ghosts synced at the end;
//   2 = carried: each stage input is synced, every padded point is advanced with edge-padded
//       stencils from it, the base state keeps its own ghost values and nothing is synced at
//       the end -- the reference's periodic step exactly (3D/src/bssn3d/evolve.py:86-90,
//       125-132: rhs() syncs a copy, the RK combinations run over the whole padded arrays).
// The interiors of 1 and 2 are identical (the same stencil values in the same order).
enum GhostMode { GH_EDGE = 0, GH_WRAP = 1, GH_CARRY = 2 };

// stage input for a carried-ghost step: Yin synced in place (a stage buffer), or for the first
// stage a synced copy of the base state in `spare`
int carried_input(const double*& yin, double* spare, size_t NS, const Dims& full,
                  hipStream_t st) {
  if (spare) {
    TRY(check_hip(hipMemcpyAsync(spare, yin, NS * sizeof(double), hipMemcpyDeviceToDevice, st),
                  "stage input copy"));
    yin = spare;
  }
  return check_hip(launch_periodic_sync(const_cast<double*>(yin), full, st), "stage sync");
}

// RK4 step; with tdev set, every stage reads its time from device memory (*tdev + offset);
// with Sdev set, the whole scalar pack is read from device memory (S16, t and dt unused):
// each stage forms dt, its time and its coefficient there, in the host's rounding.
int rk4_step(double* U, double* work, int Sx, int Sy, int Sz, const double* S16,
             const double* Sdev, int fd_order, double ko_sigma, int ng, int periodic, double t,
             const double* tdev, double dt, double chi_floor, double alpha_floor, void* stream) {
  if (!U || !work) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (periodic < 0 || periodic > 2) return fail(BSSN_E_ARG, "periodic must be 0, 1 or 2");
  if (!Sdev) {
    if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(t))
      return fail(BSSN_E_ARG, "dt must be finite and nonzero, t finite");
    TRY(check_s16(S16));
  }
  TRY(check_shape(Sx, Sy, Sz));
  if (periodic) TRY(check_ng(Sx, Sy, Sz, ng));
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)Sx * Sy * Sz;
  const size_t NS = (size_t)BSSN_NF * P;
  double* Ya = work;
  double* Yb = work + NS;
  double* KS = work + 2 * NS;
  // host-side scalars (a device pack replaces them in the kernel)
  static const double kNoPack[16] = {0, 0, 1, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1};
  const double* S = Sdev ? kNoPack : S16;
  const double h3[3] = {S[13], S[14], S[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const bool wrap = periodic == GH_WRAP, carry = periodic == GH_CARRY && ng > 0;
  const Dims full = full_dims(Sx, Sy, Sz, ng);
  const Dims d = wrap ? interior_dims(Sx, Sy, Sz, ng) : full;
  const int mode = wrap ? WRAP : CLAMP;
  const bool ko = ko_sigma > 0.0;
  StageArgs a{};
  a.U = U;
  a.K1 = KS;
  a.Uout = U;
  a.dt = dt;
  a.chi_floor = chi_floor;
  a.alpha_floor = alpha_floor;
  a.tdev = tdev;
  a.sdev = Sdev;
  a.ko_sigma = ko_sigma;
  // stage time t + tcoef dt, coefficient c0 = dt c0mul / c0div (host and device alike)
  auto stage = [&](int kind, double tcoef, double c0mul, double c0div, const double* yin,
                   double* y, double* spare) {
    if (carry) TRY(carried_input(yin, spare, NS, full, st));
    const double toff = tcoef * dt;
    a.Yin = yin; a.Y = y; a.kind = kind; a.toff = toff;
    a.c0 = dt * c0mul / c0div;
    a.tcoef = tcoef; a.c0mul = c0mul; a.c0div = c0div;
    return run_stage(a, d, w, make_scal(S, dt, t + toff), fd_order, ko, mode, st);
  };
  // k1 = f(y, t);          Y = y + (dt/2) k1    (carried: the synced copy of y lives in Yb)
  TRY(stage(KIND_RK1, 0.0, 0.5, 1.0, U, Ya, Yb));
  // k2 = f(Y, t + dt/2);   Y = y + (dt/2) k2
  TRY(stage(KIND_RK2, 0.5, 0.5, 1.0, Ya, Yb, nullptr));
  // k3 = f(Y, t + dt/2);   Y = y + dt k3
  TRY(stage(KIND_RK3, 0.5, 1.0, 1.0, Yb, Ya, nullptr));
  // k4 = f(Y, t + dt);     y + (dt/6)(k1 + 2k2 + 2k3 + k4), then enforce
  TRY(stage(KIND_RK4, 1.0, 1.0, 6.0, Ya, nullptr, nullptr));
  if (wrap && ng > 0)
    return check_hip(launch_periodic_sync(U, full, st), "periodic sync");
  return BSSN_OK;
}

}  // namespace

extern "C" {

int bssn_abi_version(void) { return BSSN_HIP_ABI_VERSION; }

const char* bssn_last_error(void) { return g_err.c_str(); }

int bssn_rhs_f64(const double* F, const double* S16, double* OUT, int Sx, int Sy, int Sz,
                 int fd_order, double ko_sigma, void* stream) {
  if (!F || !OUT) return fail(BSSN_E_ARG, "null state pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
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

int bssn_rhs_sdev_f64(const double* F, const double* S16dev, double* OUT, int Sx, int Sy, int Sz,
                      int fd_order, double ko_sigma, void* stream) {
  if (!F || !OUT || !S16dev) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  TRY(check_shape(Sx, Sy, Sz));
  const double h3[3] = {1.0, 1.0, 1.0};
  const Scales w = make_scales(h3, fd_order, ko_sigma);   // replaced from the pack
  Scal sc{};
  const Dims d = full_dims(Sx, Sy, Sz, 0);
  StageArgs a{};
  a.Yin = F;
  a.Y = OUT;
  a.kind = KIND_RHS;
  a.sdev = S16dev;
  a.ko_sigma = ko_sigma;
  a.c0mul = 0.0;
  a.c0div = 1.0;
  return run_stage(a, d, w, sc, fd_order, ko_sigma > 0.0, CLAMP, (hipStream_t)stream);
}

int bssn_stage_rhs_f64(const double* Y, double* OUT, int Sx, int Sy, int Sz, const double* S16,
                       int fd_order, double ko_sigma, int ng, int periodic, void* stream) {
  if (!Y || !OUT) return fail(BSSN_E_ARG, "null state pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
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
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  const Scales w = make_scales(h3, fd_order, 0.0);
  return check_hip(launch_deriv_bundle(F, D, full_dims(Sx, Sy, Sz, 0), w, fd_order,
                                       (hipStream_t)stream),
                   "deriv bundle launch");
}

int bssn_deriv_bundle_i8_f64(const double* F, double* D, int Sx, int Sy, int Sz, int fd_order,
                             const double* h3, int slices, void* stream) {
  if (!F || !D) return fail(BSSN_E_ARG, "null pointer");
  if (fd_order != 4)
    return fail(BSSN_E_ARG, "the INT8-split bundle supports fd_order 4 only (order-6 second "
                            "derivative numerators exceed int8)");
  if (slices < 1 || slices > 7) return fail(BSSN_E_ARG, "slices must be 1..7");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  const Scales w = make_scales(h3, 4, 0.0);
  return check_hip(launch_deriv_bundle_i8(F, D, full_dims(Sx, Sy, Sz, 0), w, slices,
                                          (hipStream_t)stream),
                   "INT8 bundle launch");
}

size_t bssn_bfp_bytes(int mant_bits, long long npts) {
  if (mant_bits < 2 || mant_bits > 48 || npts < 1) return 0;
  const size_t per = mant_bits <= 32 ? 4 : 6;
  return (size_t)BFP_HEADER_BYTES + (size_t)BSSN_ND * (size_t)npts * per;
}

int bssn_deriv_bundle_absmax_f64(const double* F, double* amax, int Sx, int Sy, int Sz,
                                 int fd_order, const double* h3, void* stream) {
  if (!F || !amax) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  const Scales w = make_scales(h3, fd_order, 0.0);
  return check_hip(launch_bundle_absmax(F, (unsigned long long*)amax, full_dims(Sx, Sy, Sz, 0),
                                        w, fd_order, (hipStream_t)stream),
                   "bundle max launch");
}

static int check_bfp(const int* exps, int mant_bits, BundleOut* o, BfpScales* sc) {
  if (!exps) return fail(BSSN_E_ARG, "null exponent array");
  if (mant_bits < 2 || mant_bits > 48) return fail(BSSN_E_ARG, "mant_bits must be 2..48");
  for (int i = 0; i < BSSN_ND; ++i) {
    if (exps[i] < -1000 || exps[i] > 1000)
      return fail(BSSN_E_ARG, "BFP exponents must lie in [-1000, 1000]");
    if (o) o->e[i] = exps[i];
    if (sc) sc->inv[i] = std::ldexp(1.0, -exps[i]);
  }
  if (o) o->lim = std::ldexp(1.0, mant_bits - 1);
  return BSSN_OK;
}

int bssn_deriv_bundle_bfp(const double* F, void* Q, const int* exps, int mant_bits, int Sx,
                          int Sy, int Sz, int fd_order, const double* h3, void* stream) {
  if (!F || !Q) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  BundleOut o{};
  BfpScales sc{};
  TRY(check_bfp(exps, mant_bits, &o, &sc));
  const Scales w = make_scales(h3, fd_order, 0.0);
  return check_hip(launch_bundle_bfp(F, Q, o, sc, full_dims(Sx, Sy, Sz, 0), w, fd_order,
                                     mant_bits > 32, (hipStream_t)stream),
                   "BFP bundle launch");
}

int bssn_algebra_bfp_f64(const double* F, const void* Q, int mant_bits, const double* S16,
                         double* OUT, long long npts, void* stream) {
  if (!F || !Q || !OUT) return fail(BSSN_E_ARG, "null pointer");
  if (npts < 1) return fail(BSSN_E_ARG, "npts must be >= 1");
  if (mant_bits < 2 || mant_bits > 48) return fail(BSSN_E_ARG, "mant_bits must be 2..48");
  TRY(check_s16(S16));
  const Scal sc = make_scal(S16, S16[2], S16[6]);
  return check_hip(launch_algebra_bfp(F, Q, OUT, (long)npts, sc, mant_bits > 32,
                                      (hipStream_t)stream),
                   "BFP algebra launch");
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
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
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
  return rk4_step(U, work, Sx, Sy, Sz, S16, nullptr, fd_order, ko_sigma, ng, periodic, t,
                  nullptr, dt, chi_floor, alpha_floor, stream);
}

int bssn_rk4_step_devtime_f64(double* U, double* work, int Sx, int Sy, int Sz,
                              const double* S16, int fd_order, double ko_sigma, int ng,
                              int periodic, double* tdev, double dt, double chi_floor,
                              double alpha_floor, void* stream) {
  if (!tdev) return fail(BSSN_E_ARG, "null device time");
  TRY(rk4_step(U, work, Sx, Sy, Sz, S16, nullptr, fd_order, ko_sigma, ng, periodic, 0.0, tdev,
               dt, chi_floor, alpha_floor, stream));
  return check_hip(launch_time_advance(tdev, dt, (hipStream_t)stream), "time advance");
}

int bssn_rk4_step_sdev_f64(double* U, double* work, int Sx, int Sy, int Sz, double* S16dev,
                           int fd_order, double ko_sigma, int ng, int periodic, double chi_floor,
                           double alpha_floor, void* stream) {
  if (!S16dev) return fail(BSSN_E_ARG, "null scalar pack");
  TRY(rk4_step(U, work, Sx, Sy, Sz, nullptr, S16dev, fd_order, ko_sigma, ng, periodic, 0.0,
               nullptr, 1.0, chi_floor, alpha_floor, stream));
  return check_hip(launch_pack_time_advance(S16dev, (hipStream_t)stream), "time advance");
}

int bssn_sommerfeld_rhs_f64(const double* F, double* OUT, int Sx, int Sy, int Sz, int fd_order,
                            const double* h3, int ng, const double* xmin3, void* stream) {
  if (!F || !OUT || !xmin3) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_h(h3));
  TRY(check_ng(Sx, Sy, Sz, ng));
  for (int q = 0; q < 3; ++q)
    if (!std::isfinite(xmin3[q])) return fail(BSSN_E_ARG, "non-finite grid origin");
  BcGeom g;
  for (int q = 0; q < 3; ++q) {
    g.xmin[q] = xmin3[q];
    g.h[q] = h3[q];
  }
  StageArgs a{};
  a.Yin = F;
  a.Y = OUT;
  a.kind = KIND_RHS;
  return check_hip(launch_sommerfeld(a, full_dims(Sx, Sy, Sz, ng), make_scales(h3, fd_order, 0.0),
                                     g, fd_order, (hipStream_t)stream),
                   "radiative boundary launch");
}

int bssn_rk4_step_sommerfeld_f64(double* U, double* work, int Sx, int Sy, int Sz,
                                 const double* S16, int fd_order, double ko_sigma, int ng,
                                 const double* xmin3, double t, double dt, double chi_floor,
                                 double alpha_floor, void* stream) {
  if (!U || !work || !xmin3) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(t))
    return fail(BSSN_E_ARG, "dt must be finite and nonzero, t finite");
  TRY(check_shape(Sx, Sy, Sz));
  TRY(check_s16(S16));
  TRY(check_ng(Sx, Sy, Sz, ng));
  const int reach = ko_sigma > 0.0 ? ko_order_of(fd_order) / 2 : fd_order / 2;
  if (ng < reach) return fail(BSSN_E_ARG, "ng must cover the reach of the interior stencils");
  for (int q = 0; q < 3; ++q)
    if (!std::isfinite(xmin3[q])) return fail(BSSN_E_ARG, "non-finite grid origin");
  hipStream_t st = (hipStream_t)stream;
  const long P = (long)Sx * Sy * Sz;
  const size_t NS = (size_t)BSSN_NF * P;
  double* Ya = work;
  double* Yb = work + NS;
  double* KS = work + 2 * NS;
  const double h3[3] = {S16[13], S16[14], S16[15]};
  const Scales w = make_scales(h3, fd_order, ko_sigma);
  const Scales w1 = make_scales(h3, fd_order, 0.0);
  const Dims din = interior_dims(Sx, Sy, Sz, ng);
  const Dims dall = full_dims(Sx, Sy, Sz, ng);
  BcGeom g;
  for (int q = 0; q < 3; ++q) {
    g.xmin[q] = xmin3[q];
    g.h[q] = h3[q];
  }
  const bool ko = ko_sigma > 0.0;
  const double hdt = 0.5 * dt;
  StageArgs a{};
  a.U = U;
  a.K1 = KS;
  a.Uout = U;
  a.dt = dt;
  a.chi_floor = chi_floor;
  a.alpha_floor = alpha_floor;
  // interiors by the fused kernel (clamped stencils read the ghost values as data), ghosts by
  // the radiative condition; both read the stage input and write disjoint points
  auto stage = [&](int kind, double c0, double toff, const double* yin, double* y) {
    a.Yin = yin;
    a.Y = y;
    a.kind = kind;
    a.c0 = c0;
    a.toff = toff;
    TRY(run_stage(a, din, w, make_scal(S16, dt, t + toff), fd_order, ko, CLAMP, st));
    return check_hip(launch_sommerfeld(a, dall, w1, g, fd_order, st), "radiative boundary");
  };
  TRY(stage(KIND_RK1, hdt, 0.0, U, Ya));
  TRY(stage(KIND_RK2, hdt, hdt, Ya, Yb));
  TRY(stage(KIND_RK3, dt, hdt, Yb, Ya));
  return stage(KIND_RK4, dt / 6.0, dt, Ya, nullptr);
}

int bssn_msrk_step_f64(double* U, double* fprev, double* fnext, double* work, int Sx, int Sy,
                       int Sz, const double* S16, int fd_order, double ko_sigma, int ng,
                       int periodic, double t, double dt, double chi_floor, double alpha_floor,
                       void* stream) {
  if (!U || !fprev || !fnext || !work) return fail(BSSN_E_ARG, "null pointer");
  if (fprev == fnext) return fail(BSSN_E_ARG, "fprev and fnext must be distinct buffers");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (periodic < 0 || periodic > 2) return fail(BSSN_E_ARG, "periodic must be 0, 1 or 2");
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
  const bool wrap = periodic == GH_WRAP, carry = periodic == GH_CARRY && ng > 0;
  const Dims full = full_dims(Sx, Sy, Sz, ng);
  const Dims d = wrap ? interior_dims(Sx, Sy, Sz, ng) : full;
  const int mode = wrap ? WRAP : CLAMP;
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
  auto stage = [&](int kind, double toff, const double* yin, double* y, double* spare) {
    if (carry) TRY(carried_input(yin, spare, NS, full, st));
    a.Yin = yin; a.Y = y; a.kind = kind;
    return run_stage(a, d, w, make_scal(S16, dt, t + toff), fd_order, ko, mode, st);
  };
  // k1 = f(y, t);            Y2 = y + dt (a20 k0 + a21 k1)   (carried: synced y in Yb)
  a.c0 = MS_A20; a.c1 = MS_A21;
  TRY(stage(KIND_MS1, 0.0, U, Ya, Yb));
  // k2 = f(Y2, t + c2 dt);   Y3 = y + dt (a30 k0 + a31 k1 + a32 k2)
  a.c0 = MS_A30; a.c1 = MS_A31; a.c2 = MS_A32;
  TRY(stage(KIND_MS2, c2 * dt, Ya, Yb, nullptr));
  // k3 = f(Y3, t + c3 dt);   y + dt (b0 k0 + b1 k1 + b2 k2 + b3 k3), then enforce
  TRY(stage(KIND_MS3, c3 * dt, Yb, nullptr, nullptr));
  if (wrap && ng > 0)
    return check_hip(launch_periodic_sync(U, full, st), "periodic sync");
  return BSSN_OK;
}

int bssn_constraints_f64(const double* F, double* HM, int Sx, int Sy, int Sz, int fd_order,
                         const double* h3, void* stream) {
  if (!F || !HM) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
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

long bssn_l2_static_capacity_doubles(void) { return BSSN_L2_SCRATCH_DOUBLES; }

int bssn_constraint_l2_static_f64(const double* F, int Sx, int Sy, int Sz, int fd_order,
                                  const double* h3, int ng, double* out2, void* stream) {
  if (!F || !out2) return fail(BSSN_E_ARG, "null pointer");
  TRY(check_ng(Sx, Sy, Sz, ng));
  const Dims d = interior_dims(Sx, Sy, Sz, ng);
  const int nblk = sumsq_blocks(d);
  const long P = (long)Sx * Sy * Sz;
  if (4 * P + 2L * nblk + 2 > BSSN_L2_SCRATCH_DOUBLES)
    return fail(BSSN_E_ARG, "grid too large for the static workspace: 4 Sx Sy Sz + 2 (Sx - 2 ng)"
                            " (Sy - 2 ng) + 2 must not exceed 10485760 doubles");
  double* ws = nullptr;
  TRY(check_hip(l2_scratch(&ws), "static workspace address"));
  double* HM = ws;
  double* partial = ws + 4 * P;
  double* total = partial + 2L * nblk;
  TRY(bssn_constraints_f64(F, HM, Sx, Sy, Sz, fd_order, h3, stream));
  hipStream_t st = (hipStream_t)stream;
  TRY(check_hip(launch_sumsq(HM, d, partial, nblk, st), "sumsq launch"));
  TRY(check_hip(launch_sum_partials(partial, nblk, total, st), "sum launch"));
  double host[2];
  TRY(check_hip(hipMemcpyAsync(host, total, sizeof(host), hipMemcpyDeviceToHost, st),
                "copy sums"));
  TRY(check_hip(hipStreamSynchronize(st), "stream synchronize"));
  const double cnt = (double)d.n[0] * d.n[1] * d.n[2];
  out2[0] = std::sqrt(host[0] / cnt);
  out2[1] = std::sqrt(host[1] / cnt);
  return BSSN_OK;
}

}  // extern "C"

// ================================================================================ batches

struct bssn_batch {
  int nblk = 0, ng = 0;
  long long pool_doubles = 0, interior_pts = 0, padded_pts = 0;
  BlockDesc* blk = nullptr;          // device block table
  BrickList full{}, inner{}, ghost{};
  int grid_full = 0, grid_inner = 0, grid_ghost = 0;   // one workgroup per brick
  int* ctr = nullptr;                // device queue counters
  int persistent = 0;                // 0, or the persistent wave count
  const long long* i8pre = nullptr;  // device: 5 prefix arrays of the INT8 bundle passes
  long long i8tot[5] = {0, 0, 0, 0, 0};
  long long bundle_doubles = 0;      // 138 arrays per block, block by block (BlockDesc::doff)
};

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

int cdiv(int n, int e) { return (n + e - 1) / e; }

// Brick lists of one batch, built on the host. A brick of the padded block holds a ghost point
// unless it lies inside [ng, S - ng) on every axis.
struct HostLists {
  std::vector<BrickItem> full, inner, ghost;
  int cb_full[BATCH_CHUNKS + 1], cb_inner[BATCH_CHUNKS + 1], cb_ghost[BATCH_CHUNKS + 1];
};

void block_bricks(int b, const int* S, int ng, std::vector<BrickItem>* full,
                  std::vector<BrickItem>* inner, std::vector<BrickItem>* ghost) {
  auto pack = [](int x, int y, int z) { return x | (y << 10) | (z << 20); };
  const int E[3] = {BSSN_BRICK_X, BSSN_BRICK_Y, BSSN_BRICK_Z};
  for (int x = 0; x < S[0]; x += E[0])
    for (int y = 0; y < S[1]; y += E[1])
      for (int z = 0; z < S[2]; z += E[2]) {
        if (full) full->push_back({b, pack(x, y, z)});
        if (ghost) {
          const int o[3] = {x, y, z};
          bool inside = true;
          for (int a = 0; a < 3; ++a)
            inside = inside && o[a] >= ng && o[a] + E[a] <= S[a] - ng;
          if (!inside) ghost->push_back({b, pack(x, y, z)});
        }
      }
  if (inner)
    for (int x = ng; x < S[0] - ng; x += E[0])
      for (int y = ng; y < S[1] - ng; y += E[1])
        for (int z = ng; z < S[2] - ng; z += E[2]) inner->push_back({b, pack(x, y, z)});
}

long list_count(int nblk, const int* S, int ng, int which) {
  long n = 0;
  for (int b = 0; b < nblk; ++b) {
    const int* s = S + 3 * b;
    if (which == 0) {
      n += (long)cdiv(s[0], BSSN_BRICK_X) * cdiv(s[1], BSSN_BRICK_Y) * cdiv(s[2], BSSN_BRICK_Z);
    } else if (which == 1) {
      n += (long)cdiv(s[0] - 2 * ng, BSSN_BRICK_X) * cdiv(s[1] - 2 * ng, BSSN_BRICK_Y) *
           cdiv(s[2] - 2 * ng, BSSN_BRICK_Z);
    } else {
      std::vector<BrickItem> g;
      block_bricks(b, s, ng, nullptr, nullptr, &g);
      n += (long)g.size();
    }
  }
  return n;
}

// Deal blocks to chunks, largest first, each to the chunk with the fewest bricks so far; the
// list is then written chunk by chunk. Returns the chunk bounds in cb.
std::vector<BrickItem> chunked(const std::vector<std::vector<BrickItem>>& per_block, int* cb) {
  std::vector<int> order(per_block.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return per_block[a].size() > per_block[b].size();
  });
  std::vector<long> load(BATCH_CHUNKS, 0);
  std::vector<std::vector<int>> members(BATCH_CHUNKS);
  for (int b : order) {
    int best = 0;
    for (int c = 1; c < BATCH_CHUNKS; ++c)
      if (load[c] < load[best]) best = c;
    load[best] += (long)per_block[b].size();
    members[best].push_back(b);
  }
  std::vector<BrickItem> out;
  cb[0] = 0;
  for (int c = 0; c < BATCH_CHUNKS; ++c) {
    std::sort(members[c].begin(), members[c].end());
    for (int b : members[c]) out.insert(out.end(), per_block[b].begin(), per_block[b].end());
    cb[c + 1] = (int)out.size();
  }
  return out;
}

int flat_grid(const int* cb) {
  int mx = 0;
  for (int c = 0; c < BATCH_CHUNKS; ++c) mx = std::max(mx, cb[c + 1] - cb[c]);
  return BATCH_CHUNKS * mx;
}

int check_blocks(int nblk, const long long* off, const int* S, const double* h, int ng,
                 long long pool_doubles) {
  if (nblk < 1) return fail(BSSN_E_ARG, "a batch needs at least one block");
  if (!off || !S || !h) return fail(BSSN_E_ARG, "null block table");
  if (ng < 0) return fail(BSSN_E_ARG, "ng must be >= 0");
  std::vector<std::pair<long long, long long>> spans;
  for (int b = 0; b < nblk; ++b) {
    const int* s = S + 3 * b;
    for (int a = 0; a < 3; ++a) {
      if (s[a] > 1024) return fail(BSSN_E_ARG, "block extents are limited to 1024 per axis");
      if (!(h[3 * b + a] > 0.0) || !std::isfinite(h[3 * b + a]))
        return fail(BSSN_E_ARG, "block spacings must be > 0");
    }
    TRY(check_shape(s[0], s[1], s[2]));
    TRY(check_ng(s[0], s[1], s[2], ng));
    const long long len = (long long)BSSN_NF * s[0] * s[1] * s[2];
    if (off[b] < 0 || off[b] + len > pool_doubles)
      return fail(BSSN_E_ARG, "block " + std::to_string(b) + " lies outside the pool");
    spans.push_back({off[b], off[b] + len});
  }
  std::sort(spans.begin(), spans.end());
  for (size_t i = 1; i < spans.size(); ++i)
    if (spans[i].first < spans[i - 1].second) return fail(BSSN_E_ARG, "blocks overlap in the pool");
  return BSSN_OK;
}

// S16 of a batch: its spacing entries are replaced per block, so only the others are checked
int check_s16_batch(const double* S16) {
  if (!S16) return fail(BSSN_E_ARG, "null scalar pack S16");
  double s[16];
  for (int i = 0; i < 16; ++i) s[i] = S16[i];
  s[3] = s[13] = s[14] = s[15] = 1.0;
  return check_s16(s);
}

int check_batch(const bssn_batch* b) {
  if (!b || !b->blk) return fail(BSSN_E_ARG, "null or uninitialised batch");
  return BSSN_OK;
}

BatchArgs batch_args(const bssn_batch* b, bool interior, const double* S16, double dt, double t,
                     double sigma) {
  BatchArgs a{};
  a.blk = b->blk;
  a.list = interior ? b->inner : b->full;
  a.ctr = b->ctr;
  a.persistent = b->persistent;
  a.interior = interior ? 1 : 0;
  a.base = make_scal(S16, dt, t);
  a.cahd_c = S16[1];
  a.dt = dt;
  a.sigma = sigma;
  return a;
}

int run_batch(const bssn_batch* b, const StageArgs& s, const BatchArgs& a, int ord, bool ko,
              int mode, hipStream_t st) {
  int grid = a.interior ? b->grid_inner : b->grid_full;
  if (b->persistent) {
    TRY(check_hip(launch_batch_reset(b->ctr, st), "batch queue reset"));
    grid = b->persistent;
  }
  if (grid <= 0) return BSSN_OK;
  return check_hip(launch_fused_batch(s, a, grid, ord, ko, mode, st), "batched stage launch");
}

}  // namespace

extern "C" {

size_t bssn_batch_plan_bytes(int nblk, const int* S, int ng) {
  if (nblk < 1 || !S || ng < 0) return 0;
  for (int b = 0; b < 3 * nblk; ++b)
    if (S[b] - 2 * ng < 1) return 0;
  size_t n = align_up(sizeof(BlockDesc) * (size_t)nblk);
  for (int w = 0; w < 3; ++w) n += align_up(sizeof(BrickItem) * (size_t)list_count(nblk, S, ng, w));
  n += align_up(sizeof(long long) * 5 * ((size_t)nblk + 1));   // INT8 bundle pass prefixes
  return n + align_up(sizeof(int) * BATCH_CHUNKS);
}

int bssn_batch_create(bssn_batch** out, void* plan, size_t plan_bytes, int nblk,
                      const long long* off, const int* S, const double* h, int ng,
                      long long pool_doubles, void* stream) {
  if (!out || !plan) return fail(BSSN_E_ARG, "null pointer");
  *out = nullptr;
  TRY(check_blocks(nblk, off, S, h, ng, pool_doubles));
  const size_t need = bssn_batch_plan_bytes(nblk, S, ng);
  if (plan_bytes < need)
    return fail(BSSN_E_ARG, "plan buffer too small: need " + std::to_string(need) + " bytes");
  std::vector<BlockDesc> desc((size_t)nblk);
  std::vector<std::vector<BrickItem>> pf((size_t)nblk), pi((size_t)nblk), pg((size_t)nblk);
  std::vector<long long> pre(5 * ((size_t)nblk + 1), 0);
  long long interior = 0, padded = 0, doff = 0;
  for (int b = 0; b < nblk; ++b) {
    const int* s = S + 3 * b;
    BlockDesc& d = desc[(size_t)b];
    d.off = (long)off[b];
    d.doff = (long)doff;
    d.moff = (long)padded;
    doff += (((long long)BSSN_ND * s[0] * s[1] * s[2]) + 63) / 64 * 64;
    for (int k = 0; k < 5; ++k)
      pre[(size_t)k * (nblk + 1) + b + 1] = pre[(size_t)k * (nblk + 1) + b] + i8_pass_tiles(k, s);
    d.ng = ng;
    for (int a = 0; a < 3; ++a) {
      d.S[a] = s[a];
      d.h[a] = h[3 * b + a];
    }
    block_bricks(b, s, ng, &pf[(size_t)b], &pi[(size_t)b], &pg[(size_t)b]);
    interior += (long long)(s[0] - 2 * ng) * (s[1] - 2 * ng) * (s[2] - 2 * ng);
    padded += (long long)s[0] * s[1] * s[2];
  }
  HostLists L;
  L.full = chunked(pf, L.cb_full);
  L.inner = chunked(pi, L.cb_inner);
  L.ghost = chunked(pg, L.cb_ghost);

  // plan image: block table | full | interior | ghost bricks | counters
  std::vector<unsigned char> img(need, 0);
  size_t pos = 0;
  auto put = [&](const void* src, size_t bytes) {
    std::memcpy(img.data() + pos, src, bytes);
    const size_t at = pos;
    pos += align_up(bytes);
    return at;
  };
  const size_t o_blk = put(desc.data(), sizeof(BlockDesc) * desc.size());
  const size_t o_full = put(L.full.data(), sizeof(BrickItem) * L.full.size());
  const size_t o_inner = put(L.inner.data(), sizeof(BrickItem) * L.inner.size());
  const size_t o_ghost = put(L.ghost.data(), sizeof(BrickItem) * L.ghost.size());
  const size_t o_pre = put(pre.data(), sizeof(long long) * pre.size());
  const size_t o_ctr = pos;
  hipStream_t st = (hipStream_t)stream;
  TRY(check_hip(hipMemcpyAsync(plan, img.data(), need, hipMemcpyHostToDevice, st), "plan upload"));
  TRY(check_hip(hipStreamSynchronize(st), "plan upload sync"));

  bssn_batch* bt = new bssn_batch();
  unsigned char* base = (unsigned char*)plan;
  bt->nblk = nblk;
  bt->ng = ng;
  bt->pool_doubles = pool_doubles;
  bt->interior_pts = interior;
  bt->padded_pts = padded;
  bt->blk = (BlockDesc*)(base + o_blk);
  bt->full.items = (const BrickItem*)(base + o_full);
  bt->inner.items = (const BrickItem*)(base + o_inner);
  bt->ghost.items = (const BrickItem*)(base + o_ghost);
  for (int c = 0; c <= BATCH_CHUNKS; ++c) {
    bt->full.cb[c] = L.cb_full[c];
    bt->inner.cb[c] = L.cb_inner[c];
    bt->ghost.cb[c] = L.cb_ghost[c];
  }
  bt->grid_full = flat_grid(L.cb_full);
  bt->grid_inner = flat_grid(L.cb_inner);
  bt->grid_ghost = flat_grid(L.cb_ghost);
  bt->ctr = (int*)(base + o_ctr);
  bt->i8pre = (const long long*)(base + o_pre);
  for (int k = 0; k < 5; ++k) bt->i8tot[k] = pre[(size_t)k * (nblk + 1) + nblk];
  bt->bundle_doubles = doff;
  *out = bt;
  return BSSN_OK;
}

void bssn_batch_destroy(bssn_batch* b) { delete b; }

int bssn_batch_set_launch(bssn_batch* b, int persistent_waves) {
  TRY(check_batch(b));
  if (persistent_waves < 0 || persistent_waves % BATCH_CHUNKS != 0 || persistent_waves > (1 << 20))
    return fail(BSSN_E_ARG, "persistent_waves must be 0 or a multiple of 8 (at most 2^20)");
  b->persistent = persistent_waves;
  return BSSN_OK;
}

long long bssn_batch_points(const bssn_batch* b, int interior) {
  if (!b) return -1;
  return interior ? b->interior_pts : b->padded_pts;
}

int bssn_batch_rhs_f64(const bssn_batch* b, const double* F, const double* S16, double* OUT,
                       int fd_order, double ko_sigma, void* stream) {
  TRY(check_batch(b));
  if (!F || !OUT) return fail(BSSN_E_ARG, "null pool pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  TRY(check_s16_batch(S16));
  StageArgs s{};
  s.Yin = F;
  s.Y = OUT;
  s.kind = KIND_RHS;
  const BatchArgs a = batch_args(b, false, S16, S16[2], S16[6], ko_sigma);
  return run_batch(b, s, a, fd_order, ko_sigma > 0.0, CLAMP, (hipStream_t)stream);
}

long long bssn_batch_bundle_doubles(const bssn_batch* b) { return b ? b->bundle_doubles : -1; }

int bssn_batch_deriv_bundle_i8_f64(const bssn_batch* b, const double* F, double* D, int fd_order,
                                   int slices, void* stream) {
  TRY(check_batch(b));
  if (!F || !D) return fail(BSSN_E_ARG, "null pointer");
  if (fd_order != 4)
    return fail(BSSN_E_ARG, "the INT8-split bundle supports fd_order 4 only (order-6 second "
                            "derivative numerators exceed int8)");
  if (slices < 1 || slices > 7) return fail(BSSN_E_ARG, "slices must be 1..7");
  return check_hip(launch_batch_bundle_i8(F, D, b->blk, b->i8pre, b->i8tot, b->nblk, slices,
                                          (hipStream_t)stream),
                   "batched INT8 bundle launch");
}

size_t bssn_batch_step_workspace_bytes(const bssn_batch* b, int integrator) {
  if (!b) return 0;
  const size_t pool = (size_t)b->pool_doubles * sizeof(double);
  return integrator == 0 ? 3 * pool : 2 * pool;
}

int bssn_batch_periodic_sync_f64(const bssn_batch* b, double* U, void* stream) {
  TRY(check_batch(b));
  if (!U) return fail(BSSN_E_ARG, "null pool pointer");
  if (b->ng == 0) return BSSN_OK;
  return check_hip(launch_batch_periodic(U, b->blk, b->ghost, b->grid_ghost, (hipStream_t)stream),
                   "batched periodic sync");
}

int bssn_batch_halo_f64(double* pool, long long pool_doubles, const bssn_halo_copy* map,
                        long long n, void* stream) {
  if (!pool || pool_doubles <= 0) return fail(BSSN_E_ARG, "null or empty pool");
  if (n < 0 || (n > 0 && !map)) return fail(BSSN_E_ARG, "bad halo map");
  return check_hip(launch_halo_copy(pool, pool_doubles, map, (long)n, (hipStream_t)stream),
                   "halo copy");
}

long long bssn_count_skipped_entries(int kind, const void* entries, long long n,
                                     long long pool_doubles) {
  if (n < 0 || (n > 0 && !entries) || pool_doubles <= 0 || kind < 0 || kind > 5) return -1;
  long long bad = 0;
  for (long long i = 0; i < n; ++i) {
    bool ok;
    if (kind == 0) ok = halo_entry_ok(((const bssn_halo_copy*)entries)[i], pool_doubles);
    else if (kind == 1) ok = prolong_entry_ok(((const bssn_prolong_point*)entries)[i], pool_doubles);
    else if (kind == 2) ok = block_job_ok(((const bssn_prolong_block*)entries)[i], pool_doubles);
    else if (kind == 3) ok = prolong_tile_ok(((const bssn_prolong_tile*)entries)[i], pool_doubles);
    else if (kind == 4)
      ok = prolong_tile_point_ok(((const bssn_prolong_tile_point*)entries)[i], pool_doubles);
    else
      ok = link_entry_ok(((const int*)entries)[2 * i], ((const int*)entries)[2 * i + 1],
                         pool_doubles);
    bad += !ok;
  }
  return bad;
}

}  // extern "C"

namespace {

// One RK4 step of a batch; bc = 2 refills ghosts with `sync` after every stage and at the end.
// hold = copy U's ghosts into the stage buffers first (a sync that may leave ghosts unwritten).
template <class Sync>
int batch_rk4(const bssn_batch* b, double* U, double* work, const double* S16, int fd_order,
              double ko_sigma, int bc, Sync sync, bool hold, double t, double dt,
              double chi_floor, double alpha_floor, double* tdev, hipStream_t st,
              const int* link = nullptr) {
  TRY(check_batch(b));
  if (!U || !work) return fail(BSSN_E_ARG, "null pointer");
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  if (!(ko_sigma >= 0.0)) return fail(BSSN_E_ARG, "ko_sigma must be >= 0");
  if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(t))
    return fail(BSSN_E_ARG, "dt must be finite and nonzero, t finite");
  TRY(check_s16_batch(S16));
  const size_t NS = (size_t)b->pool_doubles;
  double* Ya = work;
  double* Yb = work + NS;
  double* KS = work + 2 * NS;
  const bool interior = bc != 0;
  const int mode = link ? LINK : (bc == 1 ? WRAP : CLAMP);
  const bool ko = ko_sigma > 0.0;
  const double hdt = 0.5 * dt;
  StageArgs s{};
  s.U = U;
  s.K1 = KS;
  s.Uout = U;
  s.dt = dt;
  s.chi_floor = chi_floor;
  s.alpha_floor = alpha_floor;
  s.tdev = tdev;
  auto stage = [&](double toff, int kind, double c0, const double* yin, double* y) {
    BatchArgs a = batch_args(b, interior, S16, dt, t + toff, ko_sigma);
    a.link = link;
    a.pool_n = b->pool_doubles;
    s.toff = toff;
    s.Yin = yin;
    s.Y = y;
    s.kind = kind;
    s.c0 = c0;
    TRY(run_batch(b, s, a, fd_order, ko, mode, st));
    if (bc == 2 && y) TRY(sync(y));
    return BSSN_OK;
  };
  if (bc == 2 && hold && b->ng > 0)   // stage inputs start with U's ghosts
    TRY(check_hip(launch_batch_hold(U, Ya, Yb, b->blk, b->ghost, b->grid_ghost, st),
                  "ghost hold"));
  TRY(stage(0.0, KIND_RK1, hdt, U, Ya));
  TRY(stage(hdt, KIND_RK2, hdt, Ya, Yb));
  TRY(stage(hdt, KIND_RK3, dt, Yb, Ya));
  TRY(stage(dt, KIND_RK4, dt / 6.0, Ya, nullptr));
  if (bc == 1) TRY(bssn_batch_periodic_sync_f64(b, U, st));
  if (bc == 2) TRY(sync(U));
  if (tdev) TRY(check_hip(launch_time_advance(tdev, dt, st), "time advance"));
  return BSSN_OK;
}

int check_program(const bssn_ghost_program* p) {
  if (!p) return fail(BSSN_E_ARG, "null ghost program");
  if (p->nphase < 0 ||
      (p->nphase > 0 && (!p->prolong_end || !p->block_end || !p->copy_end)))
    return fail(BSSN_E_ARG, "bad ghost program phase table");
  long long pe = 0, be = 0, ce = 0, te = 0;
  for (int i = 0; i < p->nphase; ++i) {
    const long long t = p->tile_end ? p->tile_end[i] : 0;
    if (p->prolong_end[i] < pe || p->block_end[i] < be || p->copy_end[i] < ce || t < te)
      return fail(BSSN_E_ARG, "ghost program phase ends must not decrease");
    pe = p->prolong_end[i];
    be = p->block_end[i];
    ce = p->copy_end[i];
    te = t;
  }
  if ((pe > 0 && !p->prolongs) || (be > 0 && !p->blocks) || (ce > 0 && !p->copies) ||
      (te > 0 && (!p->tiles || !p->tile_points || p->n_tile_points < 0)))
    return fail(BSSN_E_ARG, "ghost program entries missing");
  return BSSN_OK;
}

int run_program(double* pool, long long pool_n, const bssn_ghost_program* p, hipStream_t st) {
  long long pb = 0, bb = 0, cb = 0, tb = 0;
  for (int i = 0; i < p->nphase; ++i) {
    const long long pe = p->prolong_end[i], be = p->block_end[i], ce = p->copy_end[i];
    const long long te = p->tile_end ? p->tile_end[i] : 0;
    if (pe > pb)
      TRY(check_hip(launch_prolong(pool, pool_n, p->prolongs + pb, (long)(pe - pb), st),
                    "prolongation"));
    if (te > tb)
      TRY(check_hip(launch_prolong_tiles(pool, pool_n, p->tiles + tb, (long)(te - tb),
                                         p->tile_points, p->n_tile_points, st),
                    "prolongation tiles"));
    tb = te;
    if (be > bb)
      TRY(check_hip(launch_prolong_blocks(pool, pool_n, p->blocks + bb, (long)(be - bb), st),
                    "block prolongation"));
    if (ce > cb)
      TRY(check_hip(launch_halo_copy(pool, pool_n, p->copies + cb, (long)(ce - cb), st),
                    "ghost copies"));
    pb = pe;
    bb = be;
    cb = ce;
  }
  return BSSN_OK;
}

}  // namespace

extern "C" {

int bssn_batch_rk4_step_f64(const bssn_batch* b, double* U, double* work, const double* S16,
                            int fd_order, double ko_sigma, int bc, const bssn_halo_copy* halo,
                            long long nhalo, double t, double dt, double chi_floor,
                            double alpha_floor, double* tdev, void* stream) {
  if (bc < 0 || bc > 2) return fail(BSSN_E_ARG, "bc must be 0 (edge), 1 (periodic) or 2 (halo)");
  if (bc == 2 && (nhalo < 0 || (nhalo > 0 && !halo)))
    return fail(BSSN_E_ARG, "bc = 2 needs a halo map");
  hipStream_t st = (hipStream_t)stream;
  auto sync = [&](double* pool) {
    return check_hip(launch_halo_copy(pool, b->pool_doubles, halo, (long)nhalo, st), "halo copy");
  };
  return batch_rk4(b, U, work, S16, fd_order, ko_sigma, bc, sync, true, t, dt, chi_floor,
                   alpha_floor, tdev, st);
}

int bssn_prolong_f64(double* pool, long long pool_doubles, const bssn_prolong_point* pts,
                     long long n, void* stream) {
  if (!pool || pool_doubles <= 0) return fail(BSSN_E_ARG, "null or empty pool");
  if (n < 0 || (n > 0 && !pts)) return fail(BSSN_E_ARG, "bad prolongation list");
  return check_hip(launch_prolong(pool, pool_doubles, pts, (long)n, (hipStream_t)stream),
                   "prolongation");
}

int bssn_prolong_blocks_f64(double* pool, long long pool_doubles, const bssn_prolong_block* jobs,
                            long long n, void* stream) {
  if (!pool || pool_doubles <= 0) return fail(BSSN_E_ARG, "null or empty pool");
  if (n < 0 || (n > 0 && !jobs)) return fail(BSSN_E_ARG, "bad prolongation job list");
  return check_hip(launch_prolong_blocks(pool, pool_doubles, jobs, (long)n, (hipStream_t)stream),
                   "block prolongation");
}

int bssn_ghost_fill_f64(double* pool, long long pool_doubles, const bssn_ghost_program* prog,
                        void* stream) {
  if (!pool || pool_doubles <= 0) return fail(BSSN_E_ARG, "null or empty pool");
  TRY(check_program(prog));
  return run_program(pool, pool_doubles, prog, (hipStream_t)stream);
}

int bssn_batch_rk4_step_prog_f64(const bssn_batch* b, double* U, double* work, const double* S16,
                                 int fd_order, double ko_sigma, const bssn_ghost_program* prog,
                                 double t, double dt, double chi_floor, double alpha_floor,
                                 double* tdev, void* stream) {
  TRY(check_program(prog));
  hipStream_t st = (hipStream_t)stream;
  auto sync = [&](double* pool) { return run_program(pool, b->pool_doubles, prog, st); };
  return batch_rk4(b, U, work, S16, fd_order, ko_sigma, 2, sync, false, t, dt, chi_floor,
                   alpha_floor, tdev, st);
}

int bssn_batch_rk4_step_link_f64(const bssn_batch* b, double* U, double* work,
                                 const double* S16, int fd_order, double ko_sigma,
                                 const int* link, long long link_points,
                                 const bssn_ghost_program* linked,
                                 const bssn_ghost_program* full, double t, double dt,
                                 double chi_floor, double alpha_floor, double* tdev,
                                 void* stream) {
  TRY(check_batch(b));
  if (!link) return fail(BSSN_E_ARG, "null link map");
  if (link_points != b->padded_pts)
    return fail(BSSN_E_ARG, "the link map needs one entry per padded point of the batch");
  TRY(check_program(linked));
  if (full) TRY(check_program(full));
  if (!valid_order(fd_order)) return fail(BSSN_E_ARG, "fd_order must be 4, 6 or 8");
  const int reach = ko_sigma > 0.0 ? std::max(fd_order, ko_order_of(fd_order)) / 2 : fd_order / 2;
  if (b->ng < reach) return fail(BSSN_E_ARG, "linked steps need ng >= the stencils' reach");
  if (b->pool_doubles >= (1LL << 29))
    return fail(BSSN_E_ARG, "linked steps need a pool below 2^29 doubles");
  hipStream_t st = (hipStream_t)stream;
  auto sync = [&](double* pool) { return run_program(pool, b->pool_doubles, linked, st); };
  TRY(batch_rk4(b, U, work, S16, fd_order, ko_sigma, 2, sync, false, t, dt, chi_floor,
                alpha_floor, tdev, st, link));
  if (full) TRY(run_program(U, b->pool_doubles, full, st));
  return BSSN_OK;
}

int bssn_batch_indicator_f64(const bssn_batch* b, const double* U, double* out, void* stream) {
  TRY(check_batch(b));
  if (!U || !out) return fail(BSSN_E_ARG, "null pointer");
  if (b->ng < 1) return fail(BSSN_E_ARG, "the indicator needs ng >= 1");
  return check_hip(launch_indicator(U, b->blk, b->nblk, out, (hipStream_t)stream), "indicator");
}

}  // extern "C"
