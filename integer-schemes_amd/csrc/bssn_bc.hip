// Radiative (Sommerfeld) outer boundary: the stage write-back of the ghost zones.
//
// The reference has no outer boundary condition for BSSN; this follows the pattern of its
// Maxwell-Chern-Simons solver (3D/src/mcs3d/main.py:230-250) for each of the 24 fields: in the
// ghost zones the right-hand side is the outgoing-wave condition
//     d_t u = -(x d_x u + y d_y u + z d_z u) / r,
// the derivative normal to the face a one-sided first difference (forward on the low face,
// backward on the high face), the tangential ones the centred order-ORD operator with edge
// padding; at edges and corners the z face wins over y, and y over x (the reference writes the
// faces in the order x, y, z). The interior is advanced by the fused kernels, whose clamped
// stencils read these ghost values as data; this kernel applies the same stage write-back
// (StageOut) to the ghost points, one thread per grid point.
#include <hip/hip_runtime.h>

#include <math.h>

#include "bssn_brick.h"
#include "bssn_kernels.h"
#include "bssn_params.h"
#include "bssn_stage.h"

namespace bssn {

template <int ORD, int KIND>
__global__ __launch_bounds__(256) void k_sommerfeld(StageArgs a, Dims d, Scales w, BcGeom g) {
  using C = Coef<ORD>;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ghost_count(d)) return;
  int c[3];
  ghost_point(d, t, c);
  const long p = ((long)c[0] * d.S[1] + c[1]) * d.S[2] + c[2];
  bool gh[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) gh[q] = c[q] < d.ng || c[q] >= d.S[q] - d.ng;
  if (!gh[0] && !gh[1] && !gh[2]) return;
  const int n = gh[2] ? 2 : (gh[1] ? 1 : 0);                  // the face whose rule applies
  const bool low = c[n] < d.ng;
  const long st[3] = {(long)d.S[1] * d.S[2], (long)d.S[2], 1L};
  double X[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) X[q] = g.xmin[q] + (double)(c[q] - d.ng) * g.h[q];
  double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
  r = r == 0.0 ? 1.0 : r;
  StageOut<KIND> out(a, d.P, p, true);
  // unrolled, so that the write-back operands and the final stage's metric are indexed by
  // constants and stay in registers (a run-time field index put them in scratch: 352-896 B)
#pragma unroll
  for (int f = 0; f < BSSN_NF; ++f) {
    const double* u = a.Yin + (long)f * d.P + p;
    double du[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q == n) {
        du[q] = low ? (u[st[q]] - u[0]) / g.h[q] : (u[0] - u[-st[q]]) / g.h[q];
      } else {
        double acc = 0.0;
#pragma unroll
        for (int k = 1; k <= C::R; ++k) {
          const int hi = c[q] + k > d.S[q] - 1 ? d.S[q] - 1 : c[q] + k;
          const int lo = c[q] - k < 0 ? 0 : c[q] - k;
          acc += C::d1(k) * (u[(long)(hi - c[q]) * st[q]] - u[(long)(lo - c[q]) * st[q]]);
        }
        du[q] = acc * w.s1[q];
      }
    }
    out.put(f, -(X[0] * du[0] + X[1] * du[1] + X[2] * du[2]) / r);
  }
  out.metric_done();
}

template <int ORD>
static hipError_t launch_bc_ord(const StageArgs& a, const Dims& d, const Scales& w,
                                const BcGeom& g, hipStream_t st) {
  const unsigned blocks = (unsigned)((ghost_count(d) + 255) / 256);
  return with_kind<true>(a.kind, [&](auto k) -> hipError_t {
    constexpr int K = decltype(k)::value;
    hipLaunchKernelGGL((k_sommerfeld<ORD, K>), dim3(blocks), dim3(256), 0, st, a, d, w, g);
    return hipGetLastError();
  });
}

hipError_t launch_sommerfeld(const StageArgs& a, const Dims& d, const Scales& w, const BcGeom& g,
                             int ord, hipStream_t st) {
  if (d.ng <= 0) return hipSuccess;
  return ord == 4   ? launch_bc_ord<4>(a, d, w, g, st)
         : ord == 6 ? launch_bc_ord<6>(a, d, w, g, st)
                    : launch_bc_ord<8>(a, d, w, g, st);
}

}  // namespace bssn
