// kernels.hip -- gfx950 kernels of the exposure path.
//
//  ccr_path_kernel<Real>   one Monte Carlo path per lane (wave64) running simulatePath
//                          (path_body.hpp): the date loop stays in registers, per-date tables
//                          come through uniform (scalar) loads, per-path look-back state (margin
//                          ring, fixing states) lives in LDS columns. Reporting-date statistics
//                          are reduced across the wavefront and written as per-wave partial sums;
//                          max(V - C, 0) is written path-major-coalesced for the PFE.
//  ccr_reduce_kernel       fixed-order sum of the per-wave partials (bit-reproducible).
//  ccr_pfe_kernel<Real>    exact order statistics per (netting set, reporting date) by MSB radix
//                          select on the IEEE bit patterns of the non-negative exposures, then
//                          the interpolation of ccr::quantile (include/ccr/core/math.hpp:39-47).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "path_body.hpp"
#include "wave.hpp"

namespace ccrhip {

__device__ inline int xcdRemap(int b, int nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous range of logical
  // workgroups so that the workgroups of one netting set share an L2.
  const int xcd = b & 7, i = b >> 3, per = nb >> 3, rem = nb & 7;
  return xcd < rem ? xcd * (per + 1) + i : rem * (per + 1) + (xcd - rem) * per + i;
}

template <typename Real, typename Sh, bool kOptions>
__global__ void __launch_bounds__(kMaxBlock) ccr_path_kernel(PathArgs<Real> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const CCR_CONST GlobalDev& g = *toConst(a.g);
  const int b = xcdRemap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  cptr<SetDev> sp = toConst(a.sets) + toConst(a.block_set)[b];
  const CCR_CONST SetDev& S = *sp;
  const int bd = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int lb = b - S.block_begin;
  const int p = lb * bd + tid;
  const bool active = p < g.num_paths;
  const int pe = active ? p : g.num_paths - 1;  // inactive lanes replay a valid path, contribute nothing
  const int waveLocal = lb * (bd >> 6) + (tid >> 6);
  double* part = a.partials + S.partial_off + static_cast<int64_t>(waveLocal) * a.max_rep * kStats;
  Real* rows = a.rows + S.rows_off;
  Real* col = reinterpret_cast<Real*>(lds_raw) + tid;
  simulatePath<Real, Sh, false, kOptions>(a, sp, pe, col, bd, [&](int rr, const double* st9, double vp) {
    if (active) rows[static_cast<int64_t>(rr) * g.num_paths + p] = static_cast<Real>(vp);
    // Without intensities E[Q] is the count of active lanes and the CVA / DVA increments are 0:
    // only the first six statistics need wavefront sums.
    constexpr int kSummed = (Sh::kExact && Sh::NC == 0) ? 6 : kStats;
#pragma unroll
    for (int k = 0; k < kSummed; ++k) {
      const double s = waveSum(active ? st9[k] : 0.0);
      if ((tid & 63) == 0) part[rr * kStats + k] = s;
    }
    if constexpr (kSummed < kStats) {
      const double lanes = static_cast<double>(__popcll(__ballot(active)));
      if ((tid & 63) == 0) {
        part[rr * kStats + 6] = lanes;
        part[rr * kStats + 7] = 0.0;
        part[rr * kStats + 8] = 0.0;
      }
    }
  });
}

// States-only pass for American Monte Carlo: the same simulation, recording the domestic state
// and the deflator of every date of the sets that hold Bermudan swaptions.
template <typename Real, typename Sh>
__global__ void __launch_bounds__(kMaxBlock) ccr_states_kernel(PathArgs<Real> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int b = xcdRemap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  cptr<SetDev> sp = toConst(a.sets) + toConst(a.block_set)[b];
  if (sp->state_off < 0) return;
  const int bd = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int p = (b - sp->block_begin) * bd + tid;
  const int paths = toConst(a.g)->num_paths;
  const int pe = p < paths ? p : paths - 1;
  Real* col = reinterpret_cast<Real*>(lds_raw) + tid;
  simulatePath<Real, Sh, true, false>(a, sp, pe, col, bd, [](int, const double*, double) {});
}

// Sum of the per-wave partials of every (set, reporting date, statistic), in wave order.
__global__ void ccr_reduce_kernel(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                                  double inv_paths, double* out) {
  const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int per = max_rep * kStats;
  if (idx >= nsets * per) return;
  const int s = idx / per, rk = idx % per;
  const double* p = partials + sets[s].partial_off + rk;
  double sum = 0.0;
  for (int w = 0; w < waves_per_set; ++w) sum += p[static_cast<int64_t>(w) * per];
  out[idx] = sum * inv_paths;
}

template <typename Real> struct KeyOf;
template <> struct KeyOf<double> { using T = unsigned long long; static constexpr int kBits = 64; };
template <> struct KeyOf<float> { using T = unsigned int; static constexpr int kBits = 32; };

template <typename Real>
__device__ inline typename KeyOf<Real>::T keyOf(Real v) {
  // Non-negative IEEE values order like their bit patterns; -0 and +0 map to key 0.
  using K = typename KeyOf<Real>::T;
  if (!(v > Real(0))) return K(0);
  if constexpr (std::is_same<Real, double>::value) return static_cast<K>(__double_as_longlong(v));
  else return static_cast<K>(__float_as_uint(v));
}
template <typename Real>
__device__ inline Real valueOf(typename KeyOf<Real>::T k) {
  if constexpr (std::is_same<Real, double>::value) return __longlong_as_double(static_cast<long long>(k));
  else return __uint_as_float(k);
}

// lo + t (hi - lo) with a rounded multiply and a rounded add, as ccr::quantile evaluates it
// (contraction into a fused multiply-add would change the last bit).
__device__ __noinline__ double lerpUnfused(double lo, double t, double hi) {
#pragma clang fp contract(off)
  const double d = hi - lo;
  const double m = t * d;
  return lo + m;
}

constexpr int kPfeThreads = 256;
constexpr int kPfeBins = 4096;   // 12-bit digits
constexpr int kPfeCand = 2048;   // candidates finished by a bitonic sort in LDS

// Finds the bin of hist[0, kPfeBins) holding rank *rank (0-based); leaves the bin in *bin and the
// rank within it in *rank. One thread scans (the histogram is small next to a row).
__device__ inline void pfeFindBin(const unsigned int* hist, long long* rank, int* bin) {
  long long r = *rank;
  int b = 0;
  for (; b < kPfeBins - 1; ++b) {
    if (r < static_cast<long long>(hist[b])) break;
    r -= hist[b];
  }
  *rank = r;
  *bin = b;
}

// One workgroup per (set, reporting date) row. Exact order statistics lo = floor(h), hi = lo + 1
// of the row's keys (h = q (P - 1), ccr::quantile): a histogram of the top 12 key bits (with the
// count of zero exposures and the smallest positive key, which settle a quantile inside the
// zeros at once), a histogram of the next 12 bits inside the selected bin, then the keys sharing
// the selected 24-bit prefix are gathered and bitonic-sorted in LDS; the next key above the
// prefix, found in the same pass, gives the upper statistic when it lies outside the bin. Bins
// with more than kPfeCand keys (ties) fall back to 8-bit radix passes over the remaining bits.
template <typename Real>
__global__ void __launch_bounds__(kPfeThreads) ccr_pfe_kernel(const SetDev* sets, const Real* rows, int max_rep,
                                                              int num_paths, double level, double* out) {
  using K = typename KeyOf<Real>::T;
  constexpr int kBits = KeyOf<Real>::kBits;
  constexpr int shiftA = kBits - 12, shiftB = kBits - 24;
  __shared__ unsigned int hist[kPfeBins];
  __shared__ K cand[kPfeCand];
  __shared__ K sPrefix, sMinPos, sMinAbove;
  __shared__ long long sRank;
  __shared__ int sBin;
  __shared__ unsigned int sCount, sZero;
  const int s = static_cast<int>(blockIdx.x) / max_rep, rr = static_cast<int>(blockIdx.x) % max_rep;
  const SetDev& S = sets[s];
  if (rr >= S.num_rep) return;
  const Real* row = rows + S.rows_off + static_cast<int64_t>(rr) * num_paths;
  const int tid = static_cast<int>(threadIdx.x);
  const double h = level * static_cast<double>(num_paths - 1);
  const long long lo = static_cast<long long>(floor(h));
  const long long hi = lo + 1 < num_paths ? lo + 1 : num_paths - 1;
  auto finish = [&](K kLo, K kHi) {
    const double vlo = static_cast<double>(valueOf<Real>(kLo)), vhi = static_cast<double>(valueOf<Real>(kHi));
    out[s * max_rep + rr] = lerpUnfused(vlo, h - static_cast<double>(lo), vhi);
  };

  // Pass A: top 12 bits, zero count, smallest positive key.
  for (int b = tid; b < kPfeBins; b += kPfeThreads) hist[b] = 0;
  if (tid == 0) {
    sZero = 0;
    sMinPos = ~K(0);
    sRank = lo;
  }
  __syncthreads();
  {
    unsigned int zeros = 0;
    K minPos = ~K(0);
    for (int i = tid; i < num_paths; i += kPfeThreads) {
      const K k = keyOf<Real>(row[i]);
      if (k == 0) ++zeros;
      else if (k < minPos) minPos = k;
      atomicAdd(&hist[k >> shiftA], 1u);
    }
    atomicAdd(&sZero, zeros);
    atomicMin(&sMinPos, minPos);
  }
  __syncthreads();
  if (lo < static_cast<long long>(sZero)) {  // the quantile lies in the zero exposures
    if (tid == 0) finish(K(0), (hi == lo || hi < static_cast<long long>(sZero)) ? K(0) : sMinPos);
    return;
  }
  if (tid == 0) pfeFindBin(hist, &sRank, &sBin);
  __syncthreads();
  const K binA = static_cast<K>(sBin);

  // Pass B: the next 12 bits inside the selected bin.
  for (int b = tid; b < kPfeBins; b += kPfeThreads) hist[b] = 0;
  __syncthreads();
  for (int i = tid; i < num_paths; i += kPfeThreads) {
    const K k = keyOf<Real>(row[i]);
    if ((k >> shiftA) == binA) atomicAdd(&hist[(k >> shiftB) & 0xFFF], 1u);
  }
  __syncthreads();
  if (tid == 0) pfeFindBin(hist, &sRank, &sBin);
  __syncthreads();
  const K prefix24 = (binA << 12) | static_cast<K>(sBin);
  const unsigned int count = hist[sBin];

  if (count <= static_cast<unsigned int>(kPfeCand)) {
    // Pass C: gather the bin's keys and the smallest key above the prefix; sort in LDS.
    if (tid == 0) {
      sCount = 0;
      sMinAbove = ~K(0);
    }
    __syncthreads();
    K minAbove = ~K(0);
    for (int i = tid; i < num_paths; i += kPfeThreads) {
      const K k = keyOf<Real>(row[i]);
      const K p = k >> shiftB;
      if (p == prefix24) cand[atomicAdd(&sCount, 1u)] = k;
      else if (p > prefix24 && k < minAbove) minAbove = k;
    }
    atomicMin(&sMinAbove, minAbove);
    __syncthreads();
    int n2 = 2;
    while (n2 < static_cast<int>(count)) n2 <<= 1;
    for (int i = static_cast<int>(count) + tid; i < n2; i += kPfeThreads) cand[i] = ~K(0);
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1)
      for (int jj = k >> 1; jj > 0; jj >>= 1) {
        for (int i = tid; i < n2; i += kPfeThreads) {
          const int ixj = i ^ jj;
          if (ixj > i) {
            const bool asc = (i & k) == 0;
            const K a = cand[i], b = cand[ixj];
            if ((a > b) == asc) {
              cand[i] = b;
              cand[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    if (tid == 0) {
      const long long r = sRank;
      const K kLo = cand[r];
      const K kHi = hi == lo ? kLo : (r + 1 < static_cast<long long>(count) ? cand[r + 1] : sMinAbove);
      finish(kLo, kHi);
    }
    return;
  }

  // Fallback for bins of ties: 8-bit radix passes over the low bits, then the upper statistic.
  if (tid == 0) sPrefix = prefix24 << shiftB;
  __syncthreads();
  unsigned int* hist8 = hist;
  K mask = ~K(0) << shiftB;
  for (int shift = shiftB - 8; shift >= 0; shift -= 8) {
    hist8[tid] = 0;
    __syncthreads();
    const K prefix = sPrefix;
    for (int i = tid; i < num_paths; i += kPfeThreads) {
      const K k = keyOf<Real>(row[i]);
      if ((k & mask) == prefix) atomicAdd(&hist8[(k >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      long long rank = sRank;
      int bin = 0;
      for (; bin < 255; ++bin) {
        if (rank < static_cast<long long>(hist8[bin])) break;
        rank -= hist8[bin];
      }
      sRank = rank;
      sPrefix = prefix | (static_cast<K>(bin) << shift);
    }
    mask |= static_cast<K>(0xFF) << shift;
    __syncthreads();
  }
  const K kLo = sPrefix;
  // Upper order statistic: equal to the lower one if enough keys are <= it, else the next key above.
  if (tid == 0) {
    sCount = 0;
    sMinAbove = ~K(0);
  }
  __syncthreads();
  unsigned int cnt = 0;
  K mn = ~K(0);
  for (int i = tid; i < num_paths; i += kPfeThreads) {
    const K k = keyOf<Real>(row[i]);
    if (k <= kLo) ++cnt;
    else if (k < mn) mn = k;
  }
  atomicAdd(&sCount, cnt);
  atomicMin(&sMinAbove, mn);
  __syncthreads();
  if (tid == 0) finish(kLo, (hi == lo || static_cast<long long>(sCount) >= hi + 1) ? kLo : sMinAbove);
}

// ---------------------------------------------------------------- launchers (host side)

template <typename Real, typename Sh>
hipError_t launchShape(const PathArgs<Real>& args, int blocks, int block, size_t lds, hipStream_t stream) {
  if constexpr (Sh::NA > 0) {
    if (args.has_options) {
      hipLaunchKernelGGL((ccr_path_kernel<Real, Sh, true>), dim3(blocks), dim3(block), lds, stream, args);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((ccr_path_kernel<Real, Sh, false>), dim3(blocks), dim3(block), lds, stream, args);
  return hipGetLastError();
}

// Picks the path-kernel specialisation for the model shape (rates, assets, credits) and for the
// presence of European options.
template <typename Real>
hipError_t launchPath(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream) {
  const size_t lds = static_cast<size_t>(args.ring + args.price_rows + args.slots) * block * sizeof(Real);
  return withShape(args.nrates, args.nassets, args.ncredits, [&](auto shape) {
    return launchShape<Real, decltype(shape)>(args, blocks, block, lds, stream);
  });
}

template <typename Real>
hipError_t launchStates(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream) {
  return withShape(args.nrates, args.nassets, args.ncredits, [&](auto shape) {
    hipLaunchKernelGGL((ccr_states_kernel<Real, decltype(shape)>), dim3(blocks), dim3(block), 0, stream, args);
    return hipGetLastError();
  });
}
template hipError_t launchStates<double>(const PathArgs<double>&, int, int, hipStream_t);
template hipError_t launchStates<float>(const PathArgs<float>&, int, int, hipStream_t);

hipError_t launchReduce(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                        double inv_paths, double* out, hipStream_t stream) {
  const int n = nsets * max_rep * kStats;
  hipLaunchKernelGGL(ccr_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, sets, partials, nsets, max_rep,
                     waves_per_set, inv_paths, out);
  return hipGetLastError();
}

template <typename Real>
hipError_t launchPfe(const SetDev* sets, const Real* rows, int nsets, int max_rep, int num_paths, double level,
                     double* out, hipStream_t stream) {
  hipLaunchKernelGGL(ccr_pfe_kernel<Real>, dim3(nsets * max_rep), dim3(256), 0, stream, sets, rows, max_rep, num_paths,
                     level, out);
  return hipGetLastError();
}

template hipError_t launchPath<double>(const PathArgs<double>&, int, int, hipStream_t);
template hipError_t launchPath<float>(const PathArgs<float>&, int, int, hipStream_t);
template hipError_t launchPfe<double>(const SetDev*, const double*, int, int, int, double, double*, hipStream_t);
template hipError_t launchPfe<float>(const SetDev*, const float*, int, int, int, double, double*, hipStream_t);

}  // namespace ccrhip
