// wave.hpp -- fixed-order wavefront and workgroup sums (DPP lane moves, no atomics), shared by
// the exposure and the American Monte Carlo kernels. Device code only.
#pragma once

#include <hip/hip_runtime.h>

namespace ccrhip {

// DPP move of a double (two 32-bit halves) with a given control; lanes of rows outside row_mask
// keep 0.
template <int Ctrl, int RowMask>
__device__ inline double dppMove(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), Ctrl, RowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), Ctrl, RowMask, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront in a fixed order (bit-reproducible), with DPP lane moves
// only: pairs (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows (row_half_mirror),
// rows (row_mirror), then row_bcast:15 and row_bcast:31 carry the row sums up; lane 63 ends
// with the total, which is broadcast.
__device__ inline double waveSum(double v) {
  v += dppMove<0xb1, 0xf>(v);
  v += dppMove<0x4e, 0xf>(v);
  v += dppMove<0x141, 0xf>(v);
  v += dppMove<0x140, 0xf>(v);
  v += dppMove<0x142, 0xa>(v);
  v += dppMove<0x143, 0xc>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Sums N values over a workgroup of `waves` wavefronts (<= 16) in a fixed order: DPP within
// each wave, then wave 0, 1, ... in order. `lds` holds waves x N doubles. The totals are valid
// in thread 0.
template <int N>
__device__ inline void blockSum(double (&v)[N], double* lds, int waves) {
  const int tid = static_cast<int>(threadIdx.x), w = tid >> 6;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double s = waveSum(v[i]);
    if ((tid & 63) == 0) lds[w * N + i] = s;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = lds[i];
      for (int k = 1; k < waves; ++k) s += lds[k * N + i];
      v[i] = s;
    }
  }
  __syncthreads();
}

}  // namespace ccrhip
