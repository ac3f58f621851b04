// Internal launch interface between the C ABI (bssn_abi.cpp) and the kernels.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

#include "bssn_params.h"

namespace bssn {

enum StageKind : int {
  KIND_RHS = 0,
  KIND_RK1 = 1, KIND_RK2 = 2, KIND_RK3 = 3, KIND_RK4 = 4,
  KIND_MS1 = 5, KIND_MS2 = 6, KIND_MS3 = 7
};

// MSRK rk4_2_2 (2 steps, 3 stages): Sanches, Brandt, Kalinani, Ji & Schnetter,
// arXiv:2603.05763, Table 1, solution (2) -- the exact fractions the reference uses
// (2D/src/mcs2d/msrk.py:47-52).
constexpr double MS_A20 = 1309.0 / 15500.0, MS_A21 = -31999.0 / 15500.0;
constexpr double MS_A30 = -241289.0 / 5880000.0, MS_A31 = 22846301.0 / 16170000.0,
                 MS_A32 = -936169.0 / 2587200.0;
constexpr double MS_B0 = -191.0 / 882.0, MS_B1 = 48241.0 / 59994.0,
                 MS_B2 = 193750.0 / 4351347.0, MS_B3 = 100000.0 / 271791.0;

struct StageArgs {
  const double* Yin;   // stencil source (state the RHS is evaluated on)
  const double* U;     // step base state y_n (pointwise)
  double* Y;           // next stage input, or the RHS itself (KIND_RHS)
  double* K1;          // RK4: running k1 + 2k2 + 2k3;  MSRK: k1 = f(y_n)
  double* FP;          // MSRK: f(y_{n-1}) on entry, partial combination after stage 2
  double* Uout;        // y_{n+1} (may alias U)
  int kind;
  double c0, c1, c2;   // per-kind coefficients
  double dt;
  double chi_floor, alpha_floor;
  const double* tdev;  // if set, the stage time is *tdev + toff
  double toff;
  // If sdev is set, the launch's scalars come from the 16-double pack in device memory
  // instead of the Scales / Scal arguments: dt = sdev[2], stage time sdev[6] + tcoef dt,
  // c0 = dt c0mul / c0div, stencil scales from sdev[13..15] and ko_sigma.
  const double* sdev;
  double ko_sigma, tcoef, c0mul, c0div;
};

// The scalars of a launch whose pack is in device memory (StageArgs::sdev): the same rounding
// as the host's unpacking of the same pack (bssn_abi.cpp), so both give identical results.
template <int ORD>
__device__ __forceinline__ void args_from_pack(StageArgs& a, Scales& w, Scal& sc) {
#pragma clang fp contract(off)
  const double* S = a.sdev;
  const double dt = S[PK_DT];
  w = scales_from_pack(S, ORD, a.ko_sigma);
  sc = scal_from_pack(S, dt, S[PK_T] + a.tcoef * dt);
  a.dt = uniform(dt);
  a.c0 = uniform(dt * a.c0mul / a.c0div);
}

// Calls fn(std::integral_constant<int, ORD>{}) with the FD order ord (4, 6 or 8) as a constant.
template <class Fn>
inline void with_order(int ord, Fn&& fn) {
  if (ord == 4)
    fn(std::integral_constant<int, 4>{});
  else if (ord == 6)
    fn(std::integral_constant<int, 6>{});
  else
    fn(std::integral_constant<int, 8>{});
}

hipError_t launch_fused(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc, int ord,
                        bool ko, int mode, hipStream_t st);
hipError_t launch_fused_o4(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                           bool ko, int mode, hipStream_t st);
hipError_t launch_fused_o6(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                           bool ko, int mode, hipStream_t st);
hipError_t launch_fused_o8(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                           bool ko, int mode, hipStream_t st);
// device-pack instances (StageArgs::sdev set): the bare RHS and RK4 kinds
hipError_t launch_fused_o4s(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st);
hipError_t launch_fused_o6s(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st);
hipError_t launch_fused_o8s(const StageArgs& a, const Dims& d, const Scales& w, const Scal& sc,
                            bool ko, int mode, hipStream_t st);

// ----------------------------------------------------------------------------- block batches
// A batch holds many blocks of different extents and spacings in one pool: block b is a
// (24, S0, S1, S2) state starting at pool offset off (doubles), laid out like a single grid.
struct BlockDesc {
  long off;       // pool offset of the block's field 0, in doubles
  int S[3];       // padded extents
  int ng;         // ghost width
  double h[3];    // spacings
  long doff;      // offset of the block's 138 derivative arrays in a batch bundle, in doubles
  long moff;      // first padded point of the block in a link map (sum of the earlier blocks' P)
};

// One brick of one block: the block index and the brick origin in the block's padded
// coordinates (10 bits per axis, x | y << 10 | z << 20).
struct BrickItem {
  int blk;
  int org;
};

// Brick lists are stored chunk-major, one chunk per XCD (blocks are dealt to chunks so that
// the chunks hold nearly equal brick counts and every block's bricks stay in one chunk).
constexpr int BATCH_CHUNKS = 8;

struct BrickList {
  const BrickItem* items;
  int cb[BATCH_CHUNKS + 1];   // chunk c = items[cb[c] .. cb[c+1])
};

struct BatchArgs {
  const BlockDesc* blk;
  BrickList list;
  int* ctr;          // BATCH_CHUNKS queue counters (zeroed before each persistent launch)
  int persistent;    // 0: one workgroup per brick; else waves pull bricks from the queues
  int interior;      // the list covers block interiors (else whole padded blocks)
  Scal base;         // shared scalars; cahd is set per block from its spacing
  double cahd_c, dt, sigma;
  const int* link;   // LINK mode: (source pool offset, source field stride) per padded point
  long long pool_n;  // LINK mode: pool length in doubles (map entries outside it are not read)
};

hipError_t launch_fused_batch(const StageArgs& a, const BatchArgs& b, int grid, int ord, bool ko,
                              int mode, hipStream_t st);
hipError_t launch_batch_o4(const StageArgs& a, const BatchArgs& b, int grid, bool ko, int mode,
                           hipStream_t st);
hipError_t launch_batch_o6(const StageArgs& a, const BatchArgs& b, int grid, bool ko, int mode,
                           hipStream_t st);
hipError_t launch_batch_o8(const StageArgs& a, const BatchArgs& b, int grid, bool ko, int mode,
                           hipStream_t st);
hipError_t launch_batch_reset(int* ctr, hipStream_t st);
hipError_t launch_batch_periodic(double* pool, const BlockDesc* blk, const BrickList& ghost,
                                 int grid, hipStream_t st);
long i8_pass_tiles(int k, const int* S);
hipError_t launch_batch_bundle_i8(const double* pool, double* D, const BlockDesc* blk,
                                  const long long* pre, const long long* total, int nblk,
                                  int nslice, hipStream_t st);
hipError_t launch_halo_copy(double* pool, long long pool_n, const void* map, long n,
                            hipStream_t st);
hipError_t launch_prolong(double* pool, long long pool_n, const void* pts, long n,
                          hipStream_t st);
hipError_t launch_prolong_tiles(double* pool, long long pool_n, const void* tiles, long n,
                                const void* tp, long long ntp, hipStream_t st);
hipError_t launch_prolong_blocks(double* pool, long long pool_n, const void* jobs, long n,
                                 hipStream_t st);
hipError_t launch_indicator(const double* pool, const BlockDesc* blk, int nblk, double* out,
                            hipStream_t st);
hipError_t launch_batch_hold(const double* U, double* Ya, double* Yb, const BlockDesc* blk,
                             const BrickList& ghost, int grid, hipStream_t st);
hipError_t launch_time_advance(double* tdev, double dt, hipStream_t st);
hipError_t launch_pack_time_advance(double* S, hipStream_t st);

// radiative outer boundary: grid origin (coordinate of the first interior point) and spacings
struct BcGeom {
  double xmin[3];
  double h[3];
};
hipError_t launch_sommerfeld(const StageArgs& a, const Dims& d, const Scales& w, const BcGeom& g,
                             int ord, hipStream_t st);

// two-pass (materialised derivative bundle) path
// BFP storage header (2^-e per derivative array), bytes
constexpr long BFP_HEADER_BYTES = 1280;

// where a bundle pass puts its values (one of these, per the pass)
struct BundleOut {
  double* D;                      // fp64 bundle
  unsigned long long* amax;       // per-array max |d| (bit patterns)
  void* Q;                        // BFP mantissas
  double lim;                     // 2^(mant_bits - 1)
  int e[BSSN_ND];                 // per-array exponents
};
struct BfpScales {
  double inv[BSSN_ND];            // 2^-e per array
};
hipError_t launch_bundle_absmax(const double* F, unsigned long long* amax, const Dims& d,
                                const Scales& w, int ord, hipStream_t st);
hipError_t launch_bundle_bfp(const double* F, void* Q, const BundleOut& o, const BfpScales& sc,
                             const Dims& d, const Scales& w, int ord, bool wide, hipStream_t st);
hipError_t launch_algebra_bfp(const double* F, const void* Q, double* OUT, long npts,
                              const Scal& sc, bool wide, hipStream_t st);
hipError_t launch_deriv_bundle(const double* F, double* D, const Dims& d, const Scales& w, int ord,
                               hipStream_t st);
hipError_t launch_algebra(const double* F, const double* D, double* OUT, long npts,
                          const Scal& sc, hipStream_t st);
// INT8-split (Ozaki scheme I) derivative bundle, FD order 4, nslice int8 digits
hipError_t launch_deriv_bundle_i8(const double* F, double* D, const Dims& d, const Scales& w,
                                  int nslice, hipStream_t st);
hipError_t launch_ko_add(const double* F, double* OUT, const Dims& d, const Scales& w, int ord,
                         int mode, hipStream_t st);

// auxiliary
hipError_t launch_periodic_sync(double* F, const Dims& d, hipStream_t st);
hipError_t launch_enforce(double* F, long P, double chi_floor, double alpha_floor,
                          hipStream_t st);
hipError_t launch_constraints(const double* F, double* HM, const Dims& d, const Scales& w, int ord,
                              hipStream_t st);
hipError_t launch_sumsq(const double* HM, const Dims& d, double* partial, int nblk,
                        hipStream_t st);
int sumsq_blocks(const Dims& d);
hipError_t launch_sum_partials(const double* partial, int nblk, double* total, hipStream_t st);
// doubles of the static workspace of bssn_constraint_l2_static_f64 (80 MiB)
#define BSSN_L2_SCRATCH_DOUBLES (10L * 1024 * 1024)
hipError_t l2_scratch(double** out);

}  // namespace bssn
