// Bounds of the caller-built entry lists (halo copies, prolongation points, whole-block
// prolongation jobs): every entry addresses the pool through offsets and strides the caller
// supplies, so each kernel checks its entry against the pool length before any access and
// skips, whole, an entry that would touch a double outside [0, pool_n). The host runs the same
// predicates in bssn_count_skipped_entries, so a caller can find such entries before a launch.
#pragma once

#include <hip/hip_runtime.h>

#include "bssn_hip.h"
#include "bssn_params.h"

namespace bssn {

// whole-block prolongation: padded extent limit of the LDS passes
constexpr int PB_WMAX = 24;
constexpr int PB_KMAX = PB_WMAX / 2 + 7;

// base + f * stride in [0, pool_n) for every field f (stride <= pool_n first, so the product
// cannot overflow for any pool that fits in memory)
__host__ __device__ inline bool fields_in_pool(long long base, long long stride,
                                               long long pool_n) {
  return base >= 0 && stride >= 0 && stride <= pool_n &&
         base + (long long)(BSSN_NF - 1) * stride < pool_n;
}

__host__ __device__ inline bool halo_entry_ok(const bssn_halo_copy& m, long long pool_n) {
  return fields_in_pool(m.dst, m.dst_stride, pool_n) &&
         fields_in_pool(m.src, m.src_stride, pool_n);
}

// the parent reads are clamped into its padded block (pS per axis), so the whole block must
// lie in the pool
__host__ __device__ inline bool prolong_entry_ok(const bssn_prolong_point& m, long long pool_n) {
  for (int a = 0; a < 3; ++a)
    if (m.pS[a] < 1 || m.pS[a] > 4096) return false;
  const long long P = (long long)m.pS[0] * m.pS[1] * m.pS[2];
  return fields_in_pool(m.dst, m.dst_stride, pool_n) && m.parent >= 0 &&
         m.parent + (long long)BSSN_NF * P <= pool_n;
}

// a prolongation tile: the parent's padded block in the pool, an aligned origin
__host__ __device__ inline bool prolong_tile_ok(const bssn_prolong_tile& T, long long pool_n) {
  for (int a = 0; a < 3; ++a) {
    if (T.pS[a] < 1 || T.pS[a] > 4096) return false;
    if ((T.origin[a] & 3) != 0) return false;
  }
  if (T.count < 0 || T.count > 4096 || T.first < 0) return false;
  const long long P = (long long)T.pS[0] * T.pS[1] * T.pS[2];
  return T.parent >= 0 && T.parent + (long long)BSSN_NF * P <= pool_n;
}

__host__ __device__ inline bool prolong_tile_point_ok(const bssn_prolong_tile_point& q,
                                                      long long pool_n) {
  return q.local >= 0 && q.local < 64 && fields_in_pool(q.dst, q.dst_stride, pool_n);
}

// a link map entry: every field of the point it names lies in the pool
__host__ __device__ inline bool link_entry_ok(int off, int stride, long long pool_n) {
  return fields_in_pool(off, stride, pool_n);
}

__host__ __device__ inline bool block_job_ok(const bssn_prolong_block& j, long long pool_n) {
  if (j.BS < 1 || j.ng < 0 || j.BS + 2 * j.ng > PB_WMAX) return false;
  for (int a = 0; a < 3; ++a)
    if (j.corner[a] != 0 && j.corner[a] != 1) return false;
  const long long W = j.BS + 2 * j.ng;
  const long long n = (long long)BSSN_NF * W * W * W;
  return j.dst >= 0 && j.dst + n <= pool_n && j.parent >= 0 && j.parent + n <= pool_n;
}

}  // namespace bssn
