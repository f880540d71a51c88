// amc.hip -- American Monte Carlo (Longstaff-Schwartz) kernels for Bermudan swaptions, FP64.
//
// Follows BermudanSwaption::valueWithExercise (src/instruments/bermudan_swaption.cpp:49-146):
//  1. exercise values: the tail swap at each exercise date, on the recorded states;
//  2. backward induction over exercise dates k (all swaptions at once): on the in-the-money paths
//     regress realised / D on x (masked moments -> standardisation and degree rules -> Gram and
//     X'y -> ridge + elimination with partial pivoting), exercise where e > 0 and e > prediction,
//     record the strategy value realised by the policy;
//  3. continuation on every valued date: the standardised design of x is shared by all swaptions
//     of a date, so the right-hand sides of every (date, basis function) against every
//     (swaption, exercise) strategy column are one GEMM  C = W S^T  with K = paths
//     (W[date, a][p] = z_p^a / D_p), followed by one small solve per (swaption, date).
// All reductions are per-workgroup DPP sums combined in a fixed order, so results are
// bit-reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"
#include "path_body.hpp"
#include "wave.hpp"

namespace ccrhip {

namespace {

constexpr int kWaves = kAmcChunk / 64;

__global__ void __launch_bounds__(256) amc_init_kernel(const BermDev* berms, int nB, int64_t paths, int16_t* exidx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nB * paths) return;
  exidx[i] = static_cast<int16_t>(berms[i / paths].ne);
}

template <typename Real>
__global__ void __launch_bounds__(kAmcChunk) amc_ev_kernel(AmcArgs a, const Real* bonds, const Real* fixed) {
  const int e = static_cast<int>(blockIdx.y);
  const int b = a.exlist[e];
  const CCR_CONST BermDev& bm = toConst(a.berms)[b];
  const CCR_CONST SetDev& S = toConst(a.sets)[bm.set];
  const int k = e - bm.ev_off;
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  const int64_t p = p0 < a.paths ? p0 : a.paths - 1;
  const int j = toConst(a.eidx)[bm.eidx_off + k];
  const double x = a.states[S.state_off + static_cast<int64_t>(j) * a.paths + p];
  const Real v = evalRange<Real>(toConst(a.evterms)[bm.ev_rec + k], static_cast<Real>(x),
                                 toConst(bonds) + 2 * static_cast<int64_t>(S.bond_off),
                                 toConst(fixed) + 4 * static_cast<int64_t>(S.fixed_off), toConst(a.poly) + S.poly_off,
                                 nullptr, 0);
  if (p0 < a.paths) a.ev[static_cast<int64_t>(e) * a.paths + p0] = static_cast<double>(v);
}

// State, deflator, exercise value and previously realised value of path p at exercise k of b.
struct ExerciseSample {
  double x, D, e, prev;
};
__device__ inline ExerciseSample exerciseSample(const AmcArgs& a, const CCR_CONST BermDev& bm, int k, int64_t p) {
  const CCR_CONST SetDev& S = toConst(a.sets)[bm.set];
  const int j = toConst(a.eidx)[bm.eidx_off + k];
  const int64_t at = S.state_off + static_cast<int64_t>(j) * a.paths + p;
  ExerciseSample s;
  s.x = a.states[at];
  s.D = a.states[a.state_len + at];
  s.e = a.ev[static_cast<int64_t>(bm.ev_off + k) * a.paths + p];
  s.prev = k + 1 < bm.ne ? a.strategy[static_cast<int64_t>(bm.ev_off + k + 1) * a.paths + p] : 0.0;
  return s;
}

// Masked moments n, sum x, sum x^2, sum y of the in-the-money paths (regression.cpp:37-46).
__global__ void __launch_bounds__(kAmcChunk) amc_mom1_kernel(AmcArgs a, int k) {
  __shared__ double lds[kWaves * kRegMom1];
  const int b = static_cast<int>(blockIdx.y);
  const CCR_CONST BermDev& bm = toConst(a.berms)[b];
  if (k >= bm.ne) return;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  double v[kRegMom1] = {0, 0, 0, 0};
  if (p < a.paths) {
    const ExerciseSample s = exerciseSample(a, bm, k, p);
    if (s.e > 0.0) {
      const double y = s.prev / s.D;
      v[0] = 1.0;
      v[1] = s.x;
      v[2] = s.x * s.x;
      v[3] = y;
    }
  }
  blockSum<kRegMom1>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < kRegMom1; ++i) a.part[(static_cast<int64_t>(b) * a.nchunks + blockIdx.x) * kRegMom1 + i] = v[i];
}

__global__ void amc_fin1_kernel(AmcArgs a, int k) {
  const int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= a.nberms) return;
  const BermDev& bm = a.berms[b];
  if (k >= bm.ne) return;
  double s[kRegMom1] = {0, 0, 0, 0};
  for (int c = 0; c < a.nchunks; ++c)
    for (int i = 0; i < kRegMom1; ++i) s[i] += a.part[(static_cast<int64_t>(b) * a.nchunks + c) * kRegMom1 + i];
  const RegParams r = regParams(s[0], s[1], s[2], bm.degree);
  double* o = a.regp + static_cast<int64_t>(b) * 8;
  o[0] = r.n;
  o[1] = r.center;
  o[2] = r.scale;
  o[3] = static_cast<double>(r.deg);
  o[4] = s[3];
}

// Gram lower triangle and X'y of the in-the-money paths (regression.cpp:8-27, 64-80).
__global__ void __launch_bounds__(kAmcChunk) amc_mom2_kernel(AmcArgs a, int k) {
  __shared__ double lds[kWaves * kRegMom2];
  const int b = static_cast<int>(blockIdx.y);
  const CCR_CONST BermDev& bm = toConst(a.berms)[b];
  if (k >= bm.ne) return;
  const double* rp = a.regp + static_cast<int64_t>(b) * 8;
  const int deg = static_cast<int>(rp[3]);
  if (deg <= 0) return;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  double v[kRegMom2];
#pragma unroll
  for (int i = 0; i < kRegMom2; ++i) v[i] = 0.0;
  if (p < a.paths) {
    const ExerciseSample s = exerciseSample(a, bm, k, p);
    if (s.e > 0.0) {
      const double y = s.prev / s.D;
      double phi[4];
      regRow((s.x - rp[1]) / rp[2], phi);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i > deg) continue;
#pragma unroll
        for (int c = 0; c <= i; ++c) v[triIndex(i, c)] = phi[i] * phi[c];
        v[10 + i] = phi[i] * y;
      }
    }
  }
  blockSum<kRegMom2>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < kRegMom2; ++i) a.part[(static_cast<int64_t>(b) * a.nchunks + blockIdx.x) * kRegMom2 + i] = v[i];
}

__global__ void amc_solve_kernel(AmcArgs a, int k) {
  const int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= a.nberms) return;
  const BermDev& bm = a.berms[b];
  if (k >= bm.ne) return;
  const double* rp = a.regp + static_cast<int64_t>(b) * 8;
  RegParams r;
  r.n = rp[0];
  r.center = rp[1];
  r.scale = rp[2];
  r.deg = static_cast<int>(rp[3]);
  double s[kRegMom2];
  for (int i = 0; i < kRegMom2; ++i) s[i] = 0.0;
  if (r.deg > 0)
    for (int c = 0; c < a.nchunks; ++c)
      for (int i = 0; i < kRegMom2; ++i) s[i] += a.part[(static_cast<int64_t>(b) * a.nchunks + c) * kRegMom2 + i];
  double coef[4];
  regSolve(r, s, s + 10, rp[4], coef);
  for (int i = 0; i < 4; ++i) a.coef[static_cast<int64_t>(b) * 4 + i] = coef[i];
}

// Exercise where e > 0 and e > prediction (bermudan_swaption.cpp:84-92); record the strategy.
__global__ void __launch_bounds__(kAmcChunk) amc_update_kernel(AmcArgs a, int k) {
  const int b = static_cast<int>(blockIdx.y);
  const CCR_CONST BermDev& bm = toConst(a.berms)[b];
  if (k >= bm.ne) return;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  if (p >= a.paths) return;
  const ExerciseSample s = exerciseSample(a, bm, k, p);
  const double* rp = a.regp + static_cast<int64_t>(b) * 8;
  const double* cf = a.coef + static_cast<int64_t>(b) * 4;
  double realised = s.prev;
  if (s.e > 0.0 && s.e > regPredict(s.x, rp[1], rp[2], static_cast<int>(rp[3]), cf)) {
    a.exidx[static_cast<int64_t>(b) * a.paths + p] = static_cast<int16_t>(k);
    realised = s.D * s.e;
  }
  a.strategy[static_cast<int64_t>(bm.ev_off + k) * a.paths + p] = realised;
}

// Continuation dates: moments of x over all paths, the shared design W, the shared Gram.
__global__ void __launch_bounds__(kAmcChunk) amc_cmom1_kernel(AmcArgs a) {
  __shared__ double lds[kWaves * 3];
  const int d = static_cast<int>(blockIdx.y);
  const CCR_CONST SetDev& S = toConst(a.sets)[a.cdates[2 * d]];
  const int j = a.cdates[2 * d + 1];
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  double v[3] = {0, 0, 0};
  if (p < a.paths) {
    const double x = a.states[S.state_off + static_cast<int64_t>(j) * a.paths + p];
    v[0] = 1.0;
    v[1] = x;
    v[2] = x * x;
  }
  blockSum<3>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < 3; ++i) a.part[(static_cast<int64_t>(d) * a.nchunks + blockIdx.x) * 3 + i] = v[i];
}

__global__ void amc_cfin1_kernel(AmcArgs a) {
  const int d = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (d >= a.ncdates) return;
  double s[3] = {0, 0, 0};
  for (int c = 0; c < a.nchunks; ++c)
    for (int i = 0; i < 3; ++i) s[i] += a.part[(static_cast<int64_t>(d) * a.nchunks + c) * 3 + i];
  for (int i = 0; i < 3; ++i) a.csum[3 * d + i] = s[i];
}

__global__ void __launch_bounds__(kAmcChunk) amc_design_kernel(AmcArgs a) {
  __shared__ double lds[kWaves * 10];
  const int d = static_cast<int>(blockIdx.y);
  const CCR_CONST SetDev& S = toConst(a.sets)[a.cdates[2 * d]];
  const int j = a.cdates[2 * d + 1];
  const RegParams r = regParams(a.csum[3 * d], a.csum[3 * d + 1], a.csum[3 * d + 2], 3);
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  double v[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) v[i] = 0.0;
  if (p < a.paths) {
    const int64_t at = S.state_off + static_cast<int64_t>(j) * a.paths + p;
    const double x = a.states[at], D = a.states[a.state_len + at];
    double phi[4];
    regRow((x - r.center) / r.scale, phi);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a.W[(static_cast<int64_t>(d) * 4 + i) * a.paths + p] = phi[i] / D;
#pragma unroll
      for (int c = 0; c <= i; ++c) v[triIndex(i, c)] = phi[i] * phi[c];
    }
  }
  blockSum<10>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < 10; ++i) a.part[(static_cast<int64_t>(d) * a.nchunks + blockIdx.x) * 10 + i] = v[i];
}

__global__ void amc_gram_kernel(AmcArgs a) {
  const int d = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (d >= a.ncdates) return;
  double s[10];
  for (int i = 0; i < 10; ++i) s[i] = 0.0;
  for (int c = 0; c < a.nchunks; ++c)
    for (int i = 0; i < 10; ++i) s[i] += a.part[(static_cast<int64_t>(d) * a.nchunks + c) * 10 + i];
  for (int i = 0; i < 10; ++i) a.gram[10 * d + i] = s[i];
}

// C[m][n] = sum_p A[m][p] B[n][p] (FP64, K = paths), 64 x 64 tile per workgroup, 4 x 4 per
// thread, K staged through LDS in chunks of 32; each output sums its K in a fixed order.
constexpr int kGT = 64, kGK = 32;
__global__ void __launch_bounds__(256) amc_gemm_nt_kernel(const double* A, const double* B, double* C, int M, int N,
                                                          int64_t K) {
  __shared__ double As[kGK][kGT + 1];
  __shared__ double Bs[kGK][kGT + 1];
  const int tx = static_cast<int>(threadIdx.x) % 16, ty = static_cast<int>(threadIdx.x) / 16;
  const int m0 = static_cast<int>(blockIdx.y) * kGT, n0 = static_cast<int>(blockIdx.x) * kGT;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = 0.0;
  for (int64_t k0 = 0; k0 < K; k0 += kGK) {
    for (int t = static_cast<int>(threadIdx.x); t < kGT * kGK; t += 256) {
      const int r = t / kGK, kk = t % kGK;
      const int64_t k = k0 + kk;
      As[kk][r] = (m0 + r < M && k < K) ? A[static_cast<int64_t>(m0 + r) * K + k] : 0.0;
      Bs[kk][r] = (n0 + r < N && k < K) ? B[static_cast<int64_t>(n0 + r) * K + k] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kGK; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        av[i] = As[kk][ty * 4 + i];
        bv[i] = Bs[kk][tx * 4 + i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = fma(av[i], bv[c], acc[i][c]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + c;
      if (m < M && n < N) C[static_cast<int64_t>(m) * N + n] = acc[i][c];
    }
}

// Continuation fit of every (Bermudan, date) record: target strategy[next exercise] / D on all
// paths (bermudan_swaption.cpp:110-116); records without one get degree -1 (prediction 0).
__global__ void amc_csolve_kernel(AmcArgs a) {
  const int r = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= a.nbdates) return;
  const BermDev& bm = a.berms[a.bdate_berm[r]];
  const BermDate& bd = a.bdates[r];
  const int d = a.bdate_cd[r];
  double* out = a.cont + static_cast<int64_t>(r) * kContStride;
  for (int i = 0; i < kContStride; ++i) out[i] = 0.0;
  out[1] = 1.0;
  out[6] = -1.0;
  if (d < 0 || bd.next_k >= bm.ne) return;
  const RegParams rp = regParams(a.csum[3 * d], a.csum[3 * d + 1], a.csum[3 * d + 2], bm.degree);
  const int col = bm.ev_off + bd.next_k;
  double xty[4];
  for (int i = 0; i < 4; ++i) xty[i] = a.C[static_cast<int64_t>(4 * d + i) * a.ncols + col];
  double coef[4];
  regSolve(rp, a.gram + 10 * d, xty, xty[0], coef);
  out[0] = rp.center;
  out[1] = rp.scale;
  for (int i = 0; i < 4; ++i) out[2 + i] = coef[i];
  out[6] = static_cast<double>(rp.deg);
}

hipError_t last() { return hipGetLastError(); }

}  // namespace

template <typename Real>
hipError_t launchAmc(const AmcArgs& a, const Real* bonds, const Real* fixed, int max_exercises, hipStream_t stream) {
  const int64_t P = a.paths;
  const dim3 blk(kAmcChunk);
  hipLaunchKernelGGL(amc_init_kernel, dim3(static_cast<unsigned>((a.nberms * P + 255) / 256)), dim3(256), 0, stream,
                     a.berms, a.nberms, P, a.exidx);
  if (hipError_t e = last()) return e;
  hipLaunchKernelGGL((amc_ev_kernel<Real>), dim3(a.nchunks, a.ncols), blk, 0, stream, a, bonds, fixed);
  if (hipError_t e = last()) return e;
  const dim3 gb(a.nchunks, a.nberms), gs((a.nberms + 255) / 256);
  for (int k = max_exercises - 1; k >= 0; --k) {
    hipLaunchKernelGGL(amc_mom1_kernel, gb, blk, 0, stream, a, k);
    hipLaunchKernelGGL(amc_fin1_kernel, gs, dim3(256), 0, stream, a, k);
    hipLaunchKernelGGL(amc_mom2_kernel, gb, blk, 0, stream, a, k);
    hipLaunchKernelGGL(amc_solve_kernel, gs, dim3(256), 0, stream, a, k);
    hipLaunchKernelGGL(amc_update_kernel, gb, blk, 0, stream, a, k);
    if (hipError_t e = last()) return e;
  }
  if (a.ncdates > 0) {
    const dim3 gd(a.nchunks, a.ncdates), gds((a.ncdates + 255) / 256);
    hipLaunchKernelGGL(amc_cmom1_kernel, gd, blk, 0, stream, a);
    hipLaunchKernelGGL(amc_cfin1_kernel, gds, dim3(256), 0, stream, a);
    hipLaunchKernelGGL(amc_design_kernel, gd, blk, 0, stream, a);
    hipLaunchKernelGGL(amc_gram_kernel, gds, dim3(256), 0, stream, a);
    const int M = 4 * a.ncdates, N = a.ncols;
    hipLaunchKernelGGL(amc_gemm_nt_kernel, dim3((N + kGT - 1) / kGT, (M + kGT - 1) / kGT), dim3(256), 0, stream, a.W,
                       a.strategy, a.C, M, N, P);
    if (hipError_t e = last()) return e;
  }
  hipLaunchKernelGGL(amc_csolve_kernel, dim3((a.nbdates + 255) / 256), dim3(256), 0, stream, a);
  return last();
}

template hipError_t launchAmc<double>(const AmcArgs&, const double*, const double*, int, hipStream_t);
template hipError_t launchAmc<float>(const AmcArgs&, const float*, const float*, int, hipStream_t);

}  // namespace ccrhip
