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

// DPP move of a double with all rows and banks enabled, for patterns whose source lanes all
// exist (quad permutations, row rotations and mirrors): no previous value is needed.
template <int Ctrl>
__device__ inline double dppFull(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), Ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), Ctrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront in a fixed order (bit-reproducible), with DPP lane moves
// only: pairs (quad_perm 1,0,3,2), quads (quad_perm 2,3,0,1), half rows (row_half_mirror),
// rows (row_mirror), then row_bcast:15 and row_bcast:31 carry the row sums up; lane 63 ends
// with the total, which is broadcast.
__device__ inline double waveSum(double v) {
  v += dppFull<0xb1>(v);
  v += dppFull<0x4e>(v);
  v += dppFull<0x141>(v);
  v += dppFull<0x140>(v);
  v += dppMove<0x142, 0xa>(v);
  v += dppMove<0x143, 0xc>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Wavefront sums of N values per lane, in a fixed order, for the cost of about one: each of
// log2(NP) transpose steps (NP = N rounded up to a power of two) exchanges half of the values a
// lane still carries with lane ^ m and keeps the other half (lanes with bit m set keep the upper
// half), so the values per lane halve while the lanes sharing a value index double; the one value
// left is then summed across the remaining lane bits. Lane L ends with the total of value index
// bitrev(L mod NP) (log2(NP) bits), which is written to *index (-1 past the padding values).
// Every lane of a value index holds the same total (xor butterflies add commutatively).
template <int N>
__device__ inline double waveSumMany(const double (&in)[N], int* index) {
  constexpr int NP = N <= 1 ? 1 : N <= 2 ? 2 : N <= 4 ? 4 : N <= 8 ? 8 : 16;
  constexpr int LB = NP == 1 ? 0 : NP == 2 ? 1 : NP == 4 ? 2 : NP == 8 ? 3 : 4;
  static_assert(N <= 16, "waveSumMany: at most 16 values");
  const int lane = static_cast<int>(threadIdx.x) & 63;
  double v[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) v[i] = i < N ? in[i] : 0.0;
#pragma unroll
  for (int l = 0; l < LB; ++l) {
    const int m = 1 << l, half = NP >> (l + 1);
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      const double keep = hi ? v[i + half] : v[i];
      const double send = hi ? v[i] : v[i + half];
      v[i] = keep + __shfl_xor(send, m);
    }
  }
  double s = v[0];
#pragma unroll
  for (int m = NP; m < 64; m <<= 1) s += __shfl_xor(s, m);
  int idx = 0;
#pragma unroll
  for (int l = 0; l < LB; ++l)
    if (lane & (1 << l)) idx |= NP >> (l + 1);
  *index = idx < N ? idx : -1;
  return s;
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
