// amc.hip -- American Monte Carlo (Longstaff-Schwartz) kernels for Bermudan swaptions.
//
// Follows BermudanSwaption::valueWithExercise (src/instruments/bermudan_swaption.cpp:49-146):
//  1. exercise values: the tail swap at each exercise date, on the recorded states;
//  2. backward induction over exercise dates k (all swaptions at once): on the in-the-money paths
//     regress realised / D on x (masked moments -> standardisation and degree rules -> Gram and
//     X'y -> ridge + elimination with partial pivoting), exercise where e > 0 and e > prediction,
//     record the strategy value realised by the policy;
//  3. continuation on every valued date: the standardised design of x (W[date, a][p] =
//     z_p^a / D_p) and its Gram are shared by all swaptions of the date; within an exercise epoch
//     every swaption regresses the same strategy column, so the right-hand sides of a tile of
//     dates x swaptions are one gathered GEMM over the paths (split in chunks, summed in chunk
//     order), followed by one small solve per (swaption, date);
//  4. valuation: per (valued date, path) the sum over the set's swaptions of position x value,
//     read by the path kernel.
// Regressions run in FP64 in both precisions. All reductions combine per-wavefront DPP sums in a
// fixed order, so results are bit-reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "launch.hpp"
#include "path_body.hpp"
#include "wave.hpp"

namespace ccrhip {

namespace {

constexpr int kWaves = kAmcChunk / 64;

// The bond-polynomial table of the plan's precision.
template <typename Real>
__device__ inline const Real* polyOf(const AmcArgs& a) {
  if constexpr (std::is_same<Real, double>::value) return a.poly;
  else return a.polyf;
}

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
  // The tail swap starts at the exercise date: no coupon is fixed yet.
  const Real v = evalRange<Real>(toConst(a.evterms)[bm.ev_rec + k], static_cast<Real>(x),
                                 toConst(bonds) + 2 * static_cast<int64_t>(S.bond_off),
                                 toConst(fixed) + 4 * static_cast<int64_t>(S.fixed_off), toConst(polyOf<Real>(a)) + S.poly_off,
                                 [](int) { return Real(0); });
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

// Sums N moments of item `item` over the path chunks in a fixed order with one wavefront: lane l
// takes chunks l, l + 64, ... in order, then a DPP wavefront sum. Valid in every lane.
template <int N>
__device__ inline void chunkSum(const double* part, int64_t item, int nchunks, double (&s)[N]) {
  const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = 0.0;
  for (int c = lane; c < nchunks; c += 64) {
    const double* q = part + (item * nchunks + c) * N;
#pragma unroll
    for (int i = 0; i < N; ++i) s[i] += q[i];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = waveSum(s[i]);
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

// One wavefront per swaption: standardisation and effective degree.
__global__ void __launch_bounds__(256) amc_fin1_kernel(AmcArgs a, int k) {
  const int b = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (b >= a.nberms) return;
  const BermDev& bm = a.berms[b];
  if (k >= bm.ne) return;
  double s[kRegMom1];
  chunkSum<kRegMom1>(a.part, b, a.nchunks, s);
  if ((threadIdx.x & 63) != 0) return;
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

// One wavefront per swaption: normal equations with the ridge, elimination with pivoting.
__global__ void __launch_bounds__(256) amc_solve_kernel(AmcArgs a, int k) {
  const int b = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
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
  if (r.deg > 0) {
    chunkSum<kRegMom2>(a.part, b, a.nchunks, s);
  } else {
    for (int i = 0; i < kRegMom2; ++i) s[i] = 0.0;
  }
  if ((threadIdx.x & 63) != 0) return;
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

// Continuation dates: moments of x over all paths and the Gram of the shared design. A
// workgroup covers kContGroup consecutive path chunks of one date (each thread sums its path of
// every chunk in order), so the grid stays small for plans with many continuation dates.
__global__ void __launch_bounds__(kAmcChunk) amc_cmom1_kernel(AmcArgs a) {
  __shared__ double lds[kWaves * 3];
  const int d = a.cuniq[blockIdx.y];
  const CCR_CONST SetDev& S = toConst(a.sets)[a.cdates[2 * d]];
  const double* xs = a.states + S.state_off + static_cast<int64_t>(a.cdates[2 * d + 1]) * a.paths;
  double v[3] = {0, 0, 0};
  for (int g = 0; g < kContGroup; ++g) {
    const int64_t p = (static_cast<int64_t>(blockIdx.x) * kContGroup + g) * kAmcChunk + threadIdx.x;
    if (p < a.paths) {
      const double x = xs[p];
      v[0] += 1.0;
      v[1] += x;
      v[2] += x * x;
    }
  }
  blockSum<3>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < 3; ++i) a.part[(static_cast<int64_t>(d) * a.ncchunks + blockIdx.x) * 3 + i] = v[i];
}

__global__ void __launch_bounds__(256) amc_cfin1_kernel(AmcArgs a) {
  const int i = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= a.ncuniq) return;
  const int d = a.cuniq[i];
  double s[3];
  chunkSum<3>(a.part, d, a.ncchunks, s);
  if ((threadIdx.x & 63) != 0) return;
  for (int i = 0; i < 3; ++i) a.csum[3 * d + i] = s[i];
}

// Gram partials of the shared design phi(z) per continuation date and path chunk (the design
// itself, W = phi(z) / D, is formed where the continuation GEMM loads it).
__global__ void __launch_bounds__(kAmcChunk) amc_design_kernel(AmcArgs a) {
  __shared__ double lds[kWaves * 10];
  const int d = a.cuniq[blockIdx.y];
  const CCR_CONST SetDev& S = toConst(a.sets)[a.cdates[2 * d]];
  const int j = a.cdates[2 * d + 1];
  const RegParams r = regParams(a.csum[3 * d], a.csum[3 * d + 1], a.csum[3 * d + 2], 3);
  const double* xs = a.states + S.state_off + static_cast<int64_t>(j) * a.paths;
  double v[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) v[i] = 0.0;
  for (int g = 0; g < kContGroup; ++g) {
    const int64_t p = (static_cast<int64_t>(blockIdx.x) * kContGroup + g) * kAmcChunk + threadIdx.x;
    if (p < a.paths) {
      double phi[4];
      regRow((xs[p] - r.center) / r.scale, phi);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int c = 0; c <= i; ++c) v[triIndex(i, c)] += phi[i] * phi[c];
    }
  }
  blockSum<10>(v, lds, kWaves);
  if (threadIdx.x == 0)
    for (int i = 0; i < 10; ++i) a.part[(static_cast<int64_t>(d) * a.ncchunks + blockIdx.x) * 10 + i] = v[i];
}

__global__ void __launch_bounds__(256) amc_gram_kernel(AmcArgs a) {
  const int i = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= a.ncuniq) return;
  const int d = a.cuniq[i];
  double s[10];
  chunkSum<10>(a.part, d, a.ncchunks, s);
  if ((threadIdx.x & 63) != 0) return;
  for (int i = 0; i < 10; ++i) a.gram[10 * d + i] = s[i];
}

// Continuation right-hand sides of one tile over one chunk of paths: rows (date dd, basis a) of
// the design W = phi(z) / D against the strategy column of each swaption in the tile. Wavefront
// w owns rows 8w .. 8w+7, lane c swaption c; paths are staged 64 at a time in LDS, the design
// rows formed there from the recorded states (x, D) of the tile's dates.
constexpr int kGStep = 64;
__global__ void __launch_bounds__(256) amc_cgemm_kernel(AmcArgs a) {
  __shared__ double Ws[4 * kGemmDates][kGStep];
  __shared__ double Ss[kGemmCols][kGStep + 1];
  const GemmTile t = a.tiles[blockIdx.x];
  const int kc = static_cast<int>(blockIdx.y);
  const int tid = static_cast<int>(threadIdx.x), w = tid >> 6, c = tid & 63;
  const int64_t kBeg = static_cast<int64_t>(kc) * kGemmK;
  const int64_t kEnd = kBeg + kGemmK < a.paths ? kBeg + kGemmK : a.paths;
  // Loader roles: design date tid / 32 (its 4 rows), paths 2 (tid % 32) and 2 (tid % 32) + 1 of
  // each step; S column tid / 4, 16 paths from (tid % 4) * 16.
  const int wd = tid >> 5, wp = (tid & 31) * 2;
  const int sc = tid >> 2, spo = (tid & 3) * 16;
  const double* xsrc = nullptr;
  RegParams rp;
  if (wd < t.nd) {
    const int d = a.tcd[t.d_off + wd], m = a.cmap[d];
    xsrc = a.states + toConst(a.sets)[a.cdates[2 * d]].state_off + static_cast<int64_t>(a.cdates[2 * d + 1]) * a.paths;
    rp = regParams(a.csum[3 * m], a.csum[3 * m + 1], a.csum[3 * m + 2], 3);
  }
  const int col = sc < t.nb ? a.gcols[t.col_off + sc] : -1;
  const double* ssrc = col >= 0 ? a.strategy + static_cast<int64_t>(col) * a.paths : nullptr;
  double acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = 0.0;
  for (int64_t k0 = kBeg; k0 < kEnd; k0 += kGStep) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int64_t k = k0 + wp + i;
      double phi[4] = {0.0, 0.0, 0.0, 0.0};
      double D = 1.0;
      if (xsrc && k < kEnd) {
        regRow((xsrc[k] - rp.center) / rp.scale, phi);
        D = xsrc[a.state_len + k];
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) Ws[4 * wd + b][wp + i] = phi[b] / D;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int64_t k = k0 + spo + i;
      Ss[sc][spo + i] = (ssrc && k < kEnd) ? ssrc[k] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < kGStep; ++kk) {
      const double s = Ss[c][kk];
#pragma unroll
      for (int i = 0; i < 8; ++i) acc[i] = fma(Ws[8 * w + i][kk], s, acc[i]);
    }
    __syncthreads();
  }
  if (c >= t.nb) return;
  const int cc = a.gcols[t.col_off + c];
  if (cc < 0) return;
  const int date_rec = a.berms[a.gberm[t.col_off + c]].date_rec;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int row = 8 * w + i, dd = row >> 2, basis = row & 3;
    if (dd >= t.nd) continue;
    const int r = date_rec + a.cdates[2 * a.tcd[t.d_off + dd] + 1];
    a.gpart[(static_cast<int64_t>(kc) * a.nbdates + r) * 4 + basis] = acc[i];
  }
}

// Narrow tiles (sets with at most kNarrowCols swaptions): up to kNarrowDates dates x nb
// swaptions. Thread (dd, b) of each half of the workgroup accumulates the 4 basis rows of date dd
// against swaption b's strategy column; the two halves take the two halves of every 64-path step
// and their sums are added in a fixed order at the end.
__global__ void __launch_bounds__(256) amc_cgemm_narrow_kernel(AmcArgs a) {
  __shared__ double Ws[4 * kNarrowDates][kGStep + 1];
  __shared__ double Ss[kNarrowCols][kGStep + 1];
  __shared__ double red[128][4];
  const GemmTile t = a.narrow[blockIdx.x];
  const int kc = static_cast<int>(blockIdx.y);
  const int tid = static_cast<int>(threadIdx.x), half = tid >> 7, dd = (tid >> 3) & 15, b = tid & 7;
  const int64_t kBeg = static_cast<int64_t>(kc) * kGemmK;
  const int64_t kEnd = kBeg + kGemmK < a.paths ? kBeg + kGemmK : a.paths;
  // Loader roles: design date tid / 16, paths 4 (tid % 16) .. +3 of each step; S column tid / 32,
  // paths 2 (tid % 32) and +1.
  const int wd = tid >> 4, wp = (tid & 15) * 4;
  const int sc = tid >> 5, spo = (tid & 31) * 2;
  const double* xsrc = nullptr;
  RegParams rp;
  if (wd < t.nd) {
    const int d = a.tcd[t.d_off + wd], m = a.cmap[d];
    xsrc = a.states + toConst(a.sets)[a.cdates[2 * d]].state_off + static_cast<int64_t>(a.cdates[2 * d + 1]) * a.paths;
    rp = regParams(a.csum[3 * m], a.csum[3 * m + 1], a.csum[3 * m + 2], 3);
  }
  const int col = sc < t.nb ? a.gcols[t.col_off + sc] : -1;
  const double* ssrc = col >= 0 ? a.strategy + static_cast<int64_t>(col) * a.paths : nullptr;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t k0 = kBeg; k0 < kEnd; k0 += kGStep) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t k = k0 + wp + i;
      double phi[4] = {0.0, 0.0, 0.0, 0.0};
      double D = 1.0;
      if (xsrc && k < kEnd) {
        regRow((xsrc[k] - rp.center) / rp.scale, phi);
        D = xsrc[a.state_len + k];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) Ws[4 * wd + c][wp + i] = phi[c] / D;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int64_t k = k0 + spo + i;
      Ss[sc][spo + i] = (ssrc && k < kEnd) ? ssrc[k] : 0.0;
    }
    __syncthreads();
    const int kk0 = half * (kGStep / 2);
#pragma unroll 8
    for (int kk = kk0; kk < kk0 + kGStep / 2; ++kk) {
      const double s = Ss[b][kk];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = fma(Ws[4 * dd + c][kk], s, acc[c]);
    }
    __syncthreads();
  }
  if (half == 1)
    for (int c = 0; c < 4; ++c) red[tid - 128][c] = acc[c];
  __syncthreads();
  if (half == 1 || dd >= t.nd || b >= t.nb) return;
  if (a.gcols[t.col_off + b] < 0) return;
  const int r = a.berms[a.gberm[t.col_off + b]].date_rec + a.cdates[2 * a.tcd[t.d_off + dd] + 1];
  for (int c = 0; c < 4; ++c) a.gpart[(static_cast<int64_t>(kc) * a.nbdates + r) * 4 + c] = acc[c] + red[tid][c];
}

// Continuation fit of every (Bermudan, date) record: target strategy[next exercise] / D on all
// paths (bermudan_swaption.cpp:110-116); the path-chunk partials of X'y are summed in chunk
// order. Records without an exercise left get degree -1 (prediction 0).
__global__ void __launch_bounds__(256) amc_csolve_kernel(AmcArgs a) {
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
  double xty[4] = {0, 0, 0, 0};
  for (int kc = 0; kc < a.nkc; ++kc)
    for (int i = 0; i < 4; ++i) xty[i] += a.gpart[(static_cast<int64_t>(kc) * a.nbdates + r) * 4 + i];
  const int m = a.cmap[d];  // the continuation date holding this date's moments and Gram
  const RegParams rp = regParams(a.csum[3 * m], a.csum[3 * m + 1], a.csum[3 * m + 2], bm.degree);
  double coef[4];
  regSolve(rp, a.gram + 10 * m, xty, xty[0], coef);
  out[0] = rp.center;
  out[1] = rp.scale;
  for (int i = 0; i < 4; ++i) out[2 + i] = i <= rp.deg ? coef[i] : 0.0;  // zeros above the degree
  out[6] = static_cast<double>(rp.deg);
}

// Per (valued date of a set with Bermudans, path): the sum over the set's swaptions of
// position x value. Workgroups of one path chunk land on one XCD (the grid's x extent is a
// multiple of 8), so the chunk's exercise indices are read from that XCD's L2.
template <typename Real, bool kSa>
__global__ void __launch_bounds__(kAmcChunk) amc_value_kernel(AmcArgs a, const Real* bonds, const Real* fixed) {
  const int chunk = static_cast<int>(blockIdx.x);
  if (chunk >= a.nchunks) return;
  const int d = static_cast<int>(blockIdx.y);
  const CCR_CONST SetDev& S = toConst(a.sets)[toConst(a.cdates)[2 * d]];
  const int j = toConst(a.cdates)[2 * d + 1];
  const int64_t p0 = static_cast<int64_t>(chunk) * kAmcChunk + threadIdx.x;
  const int64_t p = p0 < a.paths ? p0 : a.paths - 1;
  BasisView eng;
  eng.basis = a.basis;
  if constexpr (kSa) {
    eng.rrecs = toConst(a.rrecs);
    eng.rcont = toConst(a.rcont);
    eng.rfit = toConst(static_cast<const double*>(a.rfit));
    eng.saterms = toConst(a.saterms);
  }
  const Real v = bermudanSum<Real, kSa>(S, j, p, a.paths, toConst(a.cont), a.exidx,
                                   a.states, toConst(bonds) + 2 * static_cast<int64_t>(S.bond_off),
                                   toConst(fixed) + 4 * static_cast<int64_t>(S.fixed_off), toConst(polyOf<Real>(a)) + S.poly_off,
                                   toConst(a.clive), toConst(a.clive_rng)[2 * d], toConst(a.clive_rng)[2 * d + 1], eng);
  if (p0 < a.paths) a.bermv[static_cast<int64_t>(d) * a.paths + p0] = static_cast<double>(v);
}

// The state-basis valuation with kValuePaths paths per thread (bermudanSumN): a workgroup covers
// kValuePaths x kAmcChunk paths of one continuation date, thread t the paths t + i kAmcChunk.
// Four paths per thread: the 500-set book's AMC stage took 160 ms with one path per thread, 139
// with two, 129 with four and 148 with eight (136 VGPRs).
constexpr int kValuePaths = 4;
template <typename Real>
__global__ void __launch_bounds__(kAmcChunk) amc_value_paths_kernel(AmcArgs a, const Real* bonds, const Real* fixed) {
  const int chunk = static_cast<int>(blockIdx.x);
  const int64_t first = static_cast<int64_t>(chunk) * (kValuePaths * kAmcChunk) + threadIdx.x;
  if (static_cast<int64_t>(chunk) * kValuePaths * kAmcChunk >= a.paths) return;
  const int d = static_cast<int>(blockIdx.y);
  const CCR_CONST SetDev& S = toConst(a.sets)[toConst(a.cdates)[2 * d]];
  const int j = toConst(a.cdates)[2 * d + 1];
  int64_t p[kValuePaths];
#pragma unroll
  for (int i = 0; i < kValuePaths; ++i) {
    const int64_t q = first + static_cast<int64_t>(i) * kAmcChunk;
    p[i] = q < a.paths ? q : a.paths - 1;  // lanes past the last path replay it and store nothing
  }
  Real v[kValuePaths];
  bermudanSumN<Real, kValuePaths>(S, j, p, a.paths, toConst(a.cont), a.exidx, a.states,
                                  toConst(bonds) + 2 * static_cast<int64_t>(S.bond_off),
                                  toConst(fixed) + 4 * static_cast<int64_t>(S.fixed_off),
                                  toConst(polyOf<Real>(a)) + S.poly_off, toConst(a.clive), toConst(a.clive_rng)[2 * d],
                                  toConst(a.clive_rng)[2 * d + 1], v);
#pragma unroll
  for (int i = 0; i < kValuePaths; ++i) {
    const int64_t q = first + static_cast<int64_t>(i) * kAmcChunk;
    if (q < a.paths) a.bermv[static_cast<int64_t>(d) * a.paths + q] = static_cast<double>(v[i]);
  }
}

// Grid of amc_value_paths_kernel: its path chunks (a multiple of 8, so the workgroups of one chunk
// land on one XCD and read the chunk's exercise indices from its L2) x the continuation dates.
inline dim3 valuePathsGrid(const AmcArgs& a) {
  const int64_t chunks = (a.paths + kValuePaths * kAmcChunk - 1) / (kValuePaths * kAmcChunk);
  return dim3(static_cast<unsigned>((chunks + 7) / 8 * 8), static_cast<unsigned>(a.ncdates));
}

// ---------------------------------------------------------------- regression engine (basis choice)
// Records (layout.hpp RegRec) are regressed in the reference's order: exercise k of every
// swaption at once, k from the last exercise down, then every continuation record at once. Per
// batch: moments per kBasisChunk paths -> standardisation and degree per record -> Gram of
// [phi | y] per kBasisChunk paths (FP64 VALU, or FP32 matrix cores) -> chunk sums in chunk order
// and the FP64 solve per record -> the exercise decision (exercise batches).

// Record of grid row `y` of a batch: exercise k of swaption y (k >= 0), or continuation record y.
__device__ inline int basisRecord(const AmcArgs& a, int k, int y) {
  return k >= 0 ? a.rex[static_cast<int64_t>(y) * a.max_exercises + k] : a.nex_recs + y;
}

// Target of record rr on path p: deflator, exercise value (exercise records), numerator of the
// target (the value realised after this exercise, or the strategy of the next exercise) and
// whether the path enters the fit (exercise records: in the money; continuation: every path).
struct BasisTarget {
  double x, D, e, num;
  bool in;
};
__device__ inline BasisTarget basisTarget(const AmcArgs& a, const CCR_CONST RegRec& rr, const CCR_CONST BermDev& bm,
                                          int64_t p) {
  const CCR_CONST SetDev& S = toConst(a.sets)[bm.set];
  const int64_t at = S.state_off + static_cast<int64_t>(rr.j) * a.paths + p;
  BasisTarget t;
  t.x = a.states[at];
  t.D = a.states[a.state_len + at];
  if (rr.kind == 0) {
    t.e = a.ev[static_cast<int64_t>(bm.ev_off + rr.k) * a.paths + p];
    t.in = t.e > 0.0;
    t.num = rr.k + 1 < bm.ne ? a.strategy[static_cast<int64_t>(bm.ev_off + rr.k + 1) * a.paths + p] : 0.0;
  } else {
    t.e = 0.0;
    t.in = true;
    t.num = a.strategy[static_cast<int64_t>(bm.ev_off + rr.k) * a.paths + p];
  }
  return t;
}

// Regressors of record rr at state x: the state itself, or the swap rate F / A and annuity A of
// the tail (the annuity and float-leg terms valued in the plan's precision).
template <typename Real>
__device__ inline void basisFeatures(const AmcArgs& a, const CCR_CONST RegRec& rr, const CCR_CONST BermDev& bm,
                                     double x, const Real* bonds, const Real* fixed, double& u1, double& u2) {
  if (a.basis != kBasisSwapRateAnnuity) {
    u1 = x;
    u2 = 0.0;
    return;
  }
  const CCR_CONST SetDev& S = toConst(a.sets)[bm.set];
  cptr<Real> bb = toConst(bonds) + 2 * static_cast<int64_t>(S.bond_off);
  cptr<Real> ff = toConst(fixed) + 4 * static_cast<int64_t>(S.fixed_off);
  cptr<Real> pp = toConst(polyOf<Real>(a)) + S.poly_off;
  auto noFix = [](int) { return Real(0); };
  const Real xr = static_cast<Real>(x);
  const double A = static_cast<double>(evalRange<Real>(toConst(a.saterms)[rr.ann], xr, bb, ff, pp, noFix));
  const double F = static_cast<double>(evalRange<Real>(toConst(a.saterms)[rr.flt], xr, bb, ff, pp, noFix));
  u1 = F / A;
  u2 = A;
}

// Sums N partials laid out [chunk][N] from q0 over the chunks in a fixed order (lane l takes
// chunks l, l + 64, ...), then a DPP wavefront sum. Valid in every lane.
template <int N>
__device__ inline void chunkSumAt(const double* q0, int nchunks, double (&s)[N]) {
  const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = 0.0;
  for (int c = lane; c < nchunks; c += 64) {
    const double* q = q0 + static_cast<int64_t>(c) * N;
#pragma unroll
    for (int i = 0; i < N; ++i) s[i] += q[i];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = waveSum(s[i]);
}

__device__ inline double* recordPart(const AmcArgs& a, int rec) {
  return a.rpart + static_cast<int64_t>(rec) * a.nbc * kBasisGram;
}

constexpr int kBasisRounds = kBasisChunk / 256;  // paths per thread of a moment / Gram workgroup

// Moments n, sum u1, sum u1^2, sum u2, sum u2^2, sum y over the record's sample in one path chunk.
template <typename Real>
__global__ void __launch_bounds__(256) amc_rmom_kernel(AmcArgs a, const Real* bonds, const Real* fixed, int k) {
  __shared__ double lds[4 * kBasisMom];
  const int rec = basisRecord(a, k, static_cast<int>(blockIdx.y));
  if (rec < 0) return;
  const CCR_CONST RegRec& rr = toConst(a.rrecs)[rec];
  const CCR_CONST BermDev& bm = toConst(a.berms)[rr.b];
  double v[kBasisMom] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < kBasisRounds; ++r) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * kBasisChunk + r * 256 + threadIdx.x;
    if (p >= a.paths) break;
    const BasisTarget t = basisTarget(a, rr, bm, p);
    if (!t.in) continue;
    double u1, u2;
    basisFeatures<Real>(a, rr, bm, t.x, bonds, fixed, u1, u2);
    v[0] += 1.0;
    v[1] += u1;
    v[2] += u1 * u1;
    v[3] += u2;
    v[4] += u2 * u2;
    v[5] += t.num / t.D;
  }
  blockSum<kBasisMom>(v, lds, 4);
  if (threadIdx.x == 0)
    for (int i = 0; i < kBasisMom; ++i) recordPart(a, rec)[static_cast<int64_t>(blockIdx.x) * kBasisMom + i] = v[i];
}

// One wavefront per record: standardisation and effective degree (basisParams).
__global__ void __launch_bounds__(256) amc_rfin_kernel(AmcArgs a, int k, int ny) {
  const int y = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (y >= ny) return;
  const int rec = basisRecord(a, k, y);
  if (rec < 0) return;
  double s[kBasisMom];
  chunkSumAt<kBasisMom>(recordPart(a, rec), a.nbc, s);
  if ((threadIdx.x & 63) != 0) return;
  double rp[kBasisStride];
  basisParams(a.basis, a.berms[a.rrecs[rec].b].degree, s, rp);
  for (int i = 0; i < kBasisStride; ++i) a.rfit[static_cast<int64_t>(rec) * kBasisStride + i] = rp[i];
}

// Design row of a sample: the basis terms and the target in column 10.
template <typename Real>
__device__ inline bool basisSampleRow(const AmcArgs& a, const CCR_CONST RegRec& rr, const CCR_CONST BermDev& bm,
                                      const double* rp, int64_t p, const Real* bonds, const Real* fixed,
                                      double (&phi)[kBasisMax + 1]) {
  const BasisTarget t = basisTarget(a, rr, bm, p);
  if (!t.in) return false;
  double u1, u2;
  basisFeatures<Real>(a, rr, bm, t.x, bonds, fixed, u1, u2);
  basisRow(a.basis, rp, u1, u2, phi);
  phi[kBasisMax] = t.num / t.D;
  return true;
}

// Gram of [phi | y] over one path chunk in FP64: each thread sums its 16 paths, the block
// combines the wavefront sums in wavefront order.
template <typename Real>
__global__ void __launch_bounds__(256) amc_rgram_kernel(AmcArgs a, const Real* bonds, const Real* fixed, int k) {
  __shared__ double lds[4 * kBasisGram];
  const int rec = basisRecord(a, k, static_cast<int>(blockIdx.y));
  if (rec < 0) return;
  double rp[6];
  for (int i = 0; i < 6; ++i) rp[i] = toConst(static_cast<const double*>(a.rfit))[static_cast<int64_t>(rec) * kBasisStride + i];
  if (rp[5] < 1.0) return;  // degree 0 or no sample: no Gram
  const CCR_CONST RegRec& rr = toConst(a.rrecs)[rec];
  const CCR_CONST BermDev& bm = toConst(a.berms)[rr.b];
  double v[kBasisGram];
#pragma unroll
  for (int i = 0; i < kBasisGram; ++i) v[i] = 0.0;
  for (int r = 0; r < kBasisRounds; ++r) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * kBasisChunk + r * 256 + threadIdx.x;
    if (p >= a.paths) break;
    double phi[kBasisMax + 1];
    if (!basisSampleRow<Real>(a, rr, bm, rp, p, bonds, fixed, phi)) continue;
#pragma unroll
    for (int i = 0; i <= kBasisMax; ++i)
#pragma unroll
      for (int c = i; c <= kBasisMax; ++c) v[gramIndex(i, c)] += phi[i] * phi[c];
  }
  blockSum<kBasisGram>(v, lds, 4);
  if (threadIdx.x == 0)
    for (int i = 0; i < kBasisGram; ++i) recordPart(a, rec)[static_cast<int64_t>(blockIdx.x) * kBasisGram + i] = v[i];
}

// The same Gram on FP32 matrix cores (v_mfma_f32_16x16x4_f32): per round every wavefront stages
// the FP32 design rows [phi | y | 0...] of 64 paths in LDS (row = path, 16 columns) and takes
// 16 rank-4 updates G += R^T R, R the 4 x 16 block of rows 4s..4s+3: lane l supplies R[l / 16][l % 16]
// as both the A (16 x 4) and the B (4 x 16) operand. The accumulator holds G[m][n] at lane
// n + 16 (m / 4), element m % 4 (G is symmetric, so either orientation reads alike); the four
// wavefront Grams are added in FP64 in wavefront order.
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename Real>
__global__ void __launch_bounds__(256) amc_rgram32_kernel(AmcArgs a, const Real* bonds, const Real* fixed, int k) {
  __shared__ float rows[4][64][17];
  __shared__ float gw[4][16][17];
  const int rec = basisRecord(a, k, static_cast<int>(blockIdx.y));
  if (rec < 0) return;
  double rp[6];
  for (int i = 0; i < 6; ++i) rp[i] = toConst(static_cast<const double*>(a.rfit))[static_cast<int64_t>(rec) * kBasisStride + i];
  if (rp[5] < 1.0) return;
  const CCR_CONST RegRec& rr = toConst(a.rrecs)[rec];
  const CCR_CONST BermDev& bm = toConst(a.berms)[rr.b];
  const int tid = static_cast<int>(threadIdx.x), w = tid >> 6, lane = tid & 63;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int r = 0; r < kBasisRounds; ++r) {
    // Wavefront w takes paths [w * kBasisChunk / 4 + 64 r, + 64) of the chunk.
    const int64_t p = static_cast<int64_t>(blockIdx.x) * kBasisChunk + w * (kBasisChunk / 4) + r * 64 + lane;
    double phi[kBasisMax + 1];
    const bool in = p < a.paths && basisSampleRow<Real>(a, rr, bm, rp, p, bonds, fixed, phi);
#pragma unroll
    for (int c = 0; c < 16; ++c) rows[w][lane][c] = (in && c <= kBasisMax) ? static_cast<float>(phi[c]) : 0.0f;
    __syncthreads();
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) {
      const float v = rows[w][4 * s4 + (lane >> 4)][lane & 15];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(v, v, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) gw[w][4 * (lane >> 4) + i][lane & 15] = acc[i];
  __syncthreads();
  if (tid < kBasisGram) {
    int i = 0, rem = tid;
    while (rem >= kBasisMax + 1 - i) {
      rem -= kBasisMax + 1 - i;
      ++i;
    }
    const int c = i + rem;
    double s = 0.0;
    for (int q = 0; q < 4; ++q) s += static_cast<double>(gw[q][i][c]);
    recordPart(a, rec)[static_cast<int64_t>(blockIdx.x) * kBasisGram + tid] = s;
  }
}

// One wavefront per record: the chunk sums of the Gram and the FP64 solve (basisSolve).
__global__ void __launch_bounds__(256) amc_rsolve_kernel(AmcArgs a, int k, int ny) {
  const int y = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (y >= ny) return;
  const int rec = basisRecord(a, k, y);
  if (rec < 0) return;
  double* fit = a.rfit + static_cast<int64_t>(rec) * kBasisStride;
  if (fit[5] < 1.0) return;
  double g[kBasisGram];
  chunkSumAt<kBasisGram>(recordPart(a, rec), a.nbc, g);
  if ((threadIdx.x & 63) != 0) return;
  double rp[kBasisStride];
  for (int i = 0; i < kBasisStride; ++i) rp[i] = fit[i];
  basisSolve(a.basis, g, rp);
  for (int i = 6; i < kBasisStride; ++i) fit[i] = rp[i];
}

// Exercise where e > 0 and e > prediction (bermudan_swaption.cpp:84-92); record the strategy.
template <typename Real>
__global__ void __launch_bounds__(kAmcChunk) amc_rupdate_kernel(AmcArgs a, const Real* bonds, const Real* fixed, int k) {
  const int b = static_cast<int>(blockIdx.y);
  const int rec = basisRecord(a, k, b);
  if (rec < 0) return;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kAmcChunk + threadIdx.x;
  if (p >= a.paths) return;
  const CCR_CONST RegRec& rr = toConst(a.rrecs)[rec];
  const CCR_CONST BermDev& bm = toConst(a.berms)[b];
  const BasisTarget t = basisTarget(a, rr, bm, p);
  double realised = t.num;
  if (t.e > 0.0) {
    double rp[kBasisStride];
    for (int i = 0; i < kBasisStride; ++i)
      rp[i] = toConst(static_cast<const double*>(a.rfit))[static_cast<int64_t>(rec) * kBasisStride + i];
    double u1, u2;
    basisFeatures<Real>(a, rr, bm, t.x, bonds, fixed, u1, u2);
    if (t.e > basisPredict(a.basis, rp, u1, u2)) {
      a.exidx[static_cast<int64_t>(b) * a.paths + p] = static_cast<int16_t>(k);
      realised = t.D * t.e;
    }
  }
  a.strategy[static_cast<int64_t>(bm.ev_off + k) * a.paths + p] = realised;
}

// Per (swaption, date) record the continuation fit in the valuation's layout: the state basis as
// center, scale, c0..c3, degree (the swap-rate / annuity basis is read from its fit records);
// records without a valued continuation get degree -1.
__global__ void __launch_bounds__(256) amc_rcont_kernel(AmcArgs a) {
  const int r = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= a.nbdates) return;
  double* out = a.cont + static_cast<int64_t>(r) * kContStride;
  for (int i = 0; i < kContStride; ++i) out[i] = 0.0;
  out[1] = 1.0;
  out[6] = -1.0;
  const int rec = a.rcont[r];
  if (rec < 0 || a.basis != kBasisState) return;
  const double* rp = a.rfit + static_cast<int64_t>(rec) * kBasisStride;
  out[0] = rp[1];
  out[1] = rp[2];
  for (int i = 0; i < 4; ++i) out[2 + i] = i <= static_cast<int>(rp[5]) ? rp[6 + i] : 0.0;  // zeros above the degree
  out[6] = rp[5];
}

// Surrogate inputs per path: the date loop over the recorded states of a group of netting sets;
// running extremes of the domestic state and its value at each exercise date already passed (0
// before it), written on the dates of the group's rows (urow[j] >= 0). The exercise-date states
// are kept in registers (compile-time indexed).
__global__ void __launch_bounds__(256) amc_features_kernel(const double* states, int64_t state_off, int64_t paths,
                                                           const int32_t* urow, const double* utimes, int num_dates,
                                                           SurrogateExercises ex, float* feats) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  const int inputs = kSurrogateBase + ex.count;
  double hi = -1e300, lo = 1e300;
  float xe[kSurrogateMaxExercises];
#pragma unroll
  for (int m = 0; m < kSurrogateMaxExercises; ++m) xe[m] = 0.0f;
  for (int j = 0; j < num_dates; ++j) {
    const double x = states[state_off + static_cast<int64_t>(j) * paths + p];
    hi = x > hi ? x : hi;
    lo = x < lo ? x : lo;
#pragma unroll
    for (int m = 0; m < kSurrogateMaxExercises; ++m)
      if (m < ex.count && ex.date[m] == j) xe[m] = static_cast<float>(x);
    const int32_t u = urow[j];
    if (u < 0) continue;
    float* f = feats + (static_cast<int64_t>(u) * paths + p) * inputs;
    f[0] = static_cast<float>(utimes[u]);
    f[1] = static_cast<float>(x);
    f[2] = static_cast<float>(hi);
    f[3] = static_cast<float>(lo);
#pragma unroll
    for (int m = 0; m < kSurrogateMaxExercises; ++m)
      if (m < ex.count) f[kSurrogateBase + m] = xe[m];
  }
}

__global__ void zero_kernel(double* out, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
}

hipError_t last() { return hipGetLastError(); }

}  // namespace

hipError_t launchSurrogateFeatures(const double* states, int64_t state_off, int64_t paths, const int32_t* urow,
                                   const double* utimes, int num_dates, const SurrogateExercises& ex, float* feats,
                                   hipStream_t stream) {
  if (state_off < 0 || ex.count < 0 || ex.count > kSurrogateMaxExercises) return hipErrorInvalidValue;
  hipLaunchKernelGGL(amc_features_kernel, dim3(static_cast<unsigned>((paths + 255) / 256)), dim3(256), 0, stream, states,
                     state_off, paths, urow, utimes, num_dates, ex, feats);
  return last();
}

hipError_t launchZeroRows(double* out, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, out, n);
  return last();
}

template <typename Real>
hipError_t launchAmc(const AmcArgs& a, const Real* bonds, const Real* fixed, int max_exercises, hipStream_t stream) {
  const int64_t P = a.paths;
  const dim3 blk(kAmcChunk);
  hipLaunchKernelGGL(amc_init_kernel, dim3(static_cast<unsigned>((a.nberms * P + 255) / 256)), dim3(256), 0, stream,
                     a.berms, a.nberms, P, a.exidx);
  if (hipError_t e = last()) return e;
  hipLaunchKernelGGL((amc_ev_kernel<Real>), dim3(a.nchunks, a.ncols), blk, 0, stream, a, bonds, fixed);
  if (hipError_t e = last()) return e;
  // One wavefront per item in the finishing kernels.
  const dim3 gb(a.nchunks, a.nberms), gs((a.nberms + 3) / 4);
  if (a.basis >= 0) {
    // The regression engine: exercise batches, then every continuation record.
    auto batch = [&](int k, int ny) {
      const dim3 gr(a.nbc, ny), gf((ny + 3) / 4);
      hipLaunchKernelGGL((amc_rmom_kernel<Real>), gr, dim3(256), 0, stream, a, bonds, fixed, k);
      hipLaunchKernelGGL(amc_rfin_kernel, gf, dim3(256), 0, stream, a, k, ny);
      if (a.gram32)
        hipLaunchKernelGGL((amc_rgram32_kernel<Real>), gr, dim3(256), 0, stream, a, bonds, fixed, k);
      else
        hipLaunchKernelGGL((amc_rgram_kernel<Real>), gr, dim3(256), 0, stream, a, bonds, fixed, k);
      hipLaunchKernelGGL(amc_rsolve_kernel, gf, dim3(256), 0, stream, a, k, ny);
    };
    for (int k = max_exercises - 1; k >= 0; --k) {
      batch(k, a.nberms);
      hipLaunchKernelGGL((amc_rupdate_kernel<Real>), gb, blk, 0, stream, a, bonds, fixed, k);
      if (hipError_t e = last()) return e;
    }
    const int ncont = a.nrecs - a.nex_recs;
    if (ncont > 0) batch(-1, ncont);
    hipLaunchKernelGGL(amc_rcont_kernel, dim3((a.nbdates + 255) / 256), dim3(256), 0, stream, a);
    if (hipError_t e = last()) return e;
    if (a.ncdates > 0 && a.basis == kBasisSwapRateAnnuity)
      hipLaunchKernelGGL((amc_value_kernel<Real, true>), dim3((a.nchunks + 7) / 8 * 8, a.ncdates), blk, 0, stream,
                         a, bonds, fixed);
    else if (a.ncdates > 0)
      hipLaunchKernelGGL((amc_value_paths_kernel<Real>), valuePathsGrid(a), blk, 0, stream, a, bonds, fixed);
    return last();
  }
  for (int k = max_exercises - 1; k >= 0; --k) {
    hipLaunchKernelGGL(amc_mom1_kernel, gb, blk, 0, stream, a, k);
    hipLaunchKernelGGL(amc_fin1_kernel, gs, dim3(256), 0, stream, a, k);
    hipLaunchKernelGGL(amc_mom2_kernel, gb, blk, 0, stream, a, k);
    hipLaunchKernelGGL(amc_solve_kernel, gs, dim3(256), 0, stream, a, k);
    hipLaunchKernelGGL(amc_update_kernel, gb, blk, 0, stream, a, k);
    if (hipError_t e = last()) return e;
  }
  if (a.ncdates > 0) {
    const dim3 gd(a.ncchunks, a.ncuniq), gds((a.ncuniq + 3) / 4);
    hipLaunchKernelGGL(amc_cmom1_kernel, gd, blk, 0, stream, a);
    hipLaunchKernelGGL(amc_cfin1_kernel, gds, dim3(256), 0, stream, a);
    hipLaunchKernelGGL(amc_design_kernel, gd, blk, 0, stream, a);
    hipLaunchKernelGGL(amc_gram_kernel, gds, dim3(256), 0, stream, a);
    if (a.ntiles > 0)
      hipLaunchKernelGGL(amc_cgemm_kernel, dim3(a.ntiles, a.nkc), dim3(256), 0, stream, a);
    if (a.nnarrow > 0)
      hipLaunchKernelGGL(amc_cgemm_narrow_kernel, dim3(a.nnarrow, a.nkc), dim3(256), 0, stream, a);
    if (hipError_t e = last()) return e;
  }
  hipLaunchKernelGGL(amc_csolve_kernel, dim3((a.nbdates + 255) / 256), dim3(256), 0, stream, a);
  if (a.ncdates > 0)
    hipLaunchKernelGGL((amc_value_paths_kernel<Real>), valuePathsGrid(a), blk, 0, stream, a, bonds, fixed);
  return last();
}

template hipError_t launchAmc<double>(const AmcArgs&, const double*, const double*, int, hipStream_t);
template hipError_t launchAmc<float>(const AmcArgs&, const float*, const float*, int, hipStream_t);

}  // namespace ccrhip
