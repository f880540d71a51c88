// kernels.hip -- gfx950 kernels of the exposure path.
//
//  ccr_path_kernel<Real>   one Monte Carlo path per lane (wave64) running simulatePath
//                          (path_body.hpp): the date loop stays in registers, per-date tables
//                          come through uniform (scalar) loads, per-path look-back state (margin
//                          ring, fixing states) lives in LDS columns. Reporting-date statistics
//                          are reduced across the wavefront and written as per-wave partial sums;
//                          max(V - C, 0) is written path-major-coalesced for the PFE.
//  ccr_reduce_kernel       fixed-order sum of the per-wave partials (bit-reproducible across runs
//                          and workgroup sizes).
//  ccr_pfe_kernel<Real>    exact order statistics per (netting set, reporting date) on the IEEE
//                          bit patterns of the non-negative exposures: splitters from a sorted
//                          sample, one counting / gathering pass, an LDS sort of the keys between
//                          the splitters (MSB radix select as the fallback), then the
//                          interpolation of ccr::quantile (include/ccr/core/math.hpp:39-47).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <vector>

#include "path_body.hpp"
#include "wave.hpp"

namespace ccrhip {

__device__ inline int xcdRemap(int b, int nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous range of logical
  // workgroups so that the workgroups of one netting set share an L2.
  const int xcd = b & 7, i = b >> 3, per = nb >> 3, rem = nb & 7;
  return xcd < rem ? xcd * (per + 1) + i : rem * (per + 1) + (xcd - rem) * per + i;
}

// Minimum waves per SIMD the register allocator is held to, per specialisation: 5 (at most 96
// VGPRs) for three rates and two assets with at most one credit and no options in FP64 (97 and
// 126 VGPRs unconstrained; the few spills cost less than the fifth wave gains: -6% on the
// batched 500-set portfolio), else no constraint.
template <typename Real, typename Sh, bool kOptions>
constexpr int pathWaveHint() {
  return (std::is_same<Real, double>::value && Sh::kExact && Sh::NR == 3 && Sh::NA == 2 && Sh::NC <= 1 && !kOptions)
             ? 5 : 1;
}

// The wavefront sums of a reporting date's kStats statistics into this wave's partials (out: the
// date's kStats slots): the first eight in one transposed wavefront sum (lanes 0..7 store one
// each), the DVA increment -- 0 without intensities (Q = 1) -- in a sum of its own (stored by
// lane 8 either way). Inactive lanes (past the last path) add zeros.
template <typename Sh>
__device__ inline void storeStats(const double* st9, bool active, double* out) {
  double vals[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) vals[k] = active ? st9[k] : 0.0;
  int idx;
  const double s = waveSumMany<8>(vals, &idx);
  double dva = 0.0;
  if constexpr (!(Sh::kExact && Sh::NC == 0)) dva = waveSum(active ? st9[8] : 0.0);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  if (lane < 9) out[lane < 8 ? idx : 8] = lane < 8 ? s : dva;
}

template <typename Real, typename Sh, bool kOptions, bool kTrades = false>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(pathWaveHint<Real, Sh, kOptions>())))
ccr_path_kernel(PathArgs<Real> a) {
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
  // Initial-margin proxy (uniform switch): |V_j - V_j-1| per path and the wavefront's deflator
  // sum on every valued date.
  Real vPrev = 0;
  auto onValue = [&](int j, Real V, Real D) {
    if (!a.im_rows) return;
    if (j > 0 && active) a.im_rows[S.im_off + static_cast<int64_t>(j - 1) * g.num_paths + p] = V > vPrev ? V - vPrev : vPrev - V;
    vPrev = V;
    const double sD = waveSum(active ? static_cast<double>(D) : 0.0);
    if ((tid & 63) == 0) a.im_part[S.im_part_off + static_cast<int64_t>(waveLocal) * a.im_dates + j] = sD;
  };
  // Per-trade statistics (kTrades): on every reporting date, per trade i, D V_i 1{V>0},
  // V_i 1{V>0}, D (V+ - (V - V_i)+), D V_i+ and V_i+ summed over the wavefront (one transposed
  // sum per trade) into this wave's partials.
  auto onTrades = [&](int rr, Real V, Real D, const Real* tv, int stride) {
    double* tp = a.tpart + S.tpart_off + (static_cast<int64_t>(waveLocal) * a.max_rep + rr) * a.tstride;
    const double v = static_cast<double>(V), d = static_cast<double>(D);
    const double on = v > 0.0 ? 1.0 : 0.0, vp = v > 0.0 ? v : 0.0;
    for (int i = 0; i < S.num_groups; ++i) {
      const double vi = static_cast<double>(tv[i * stride]);
      const double rest = v - vi, vip = vi > 0.0 ? vi : 0.0;
      double vals[kTradeStats];
      vals[0] = active ? d * vi * on : 0.0;
      vals[1] = active ? vi * on : 0.0;
      vals[2] = active ? d * (vp - (rest > 0.0 ? rest : 0.0)) : 0.0;
      vals[3] = active ? d * vip : 0.0;
      vals[4] = active ? vip : 0.0;
      int idx;
      const double s = waveSumMany<kTradeStats>(vals, &idx);
      if ((tid & 63) < 8 && idx >= 0) tp[i * kTradeStats + idx] = s;
    }
  };
  simulatePath<Real, Sh, kOptions, kTrades>(a, sp, pe, col, bd, [&](int rr, const double* st9, double vp) {
    if (active) rows[static_cast<int64_t>(rr) * g.num_paths + p] = static_cast<Real>(vp);
    storeStats<Sh>(st9, active, part + rr * kStats);
  }, onValue, onTrades);
}

// Book plans: G netting sets per workgroup (simulateGroup), one path per lane. Workgroup b (XCD
// remapped) covers path block b / ngroups of the sets [G (b % ngroups), +G): path-block-major, so
// the groups reading one path block's replayed normals (every set of a book replays the same)
// run together on one XCD and share its L2. Each set's statistics go to its own wave partials and
// exposure rows as in ccr_path_kernel. LDS per path: the shared fixing slots, then each set's
// margin-call ring and price rows, then (rolled set loops) its state rows.
// The book kernel's registers are held to five waves per SIMD (96 VGPRs). The unrolled four-set
// loop keeps each set's state (collateral units, previous V - C: two values per set) in
// registers and fits that with a few bytes of spills in FP64 (FP32: 88 VGPRs, none); the
// 500-set book's FP64 path kernel took 317 ms at five waves, 320 ms at six (80 VGPRs, 76 bytes
// of spills per lane) and 322 ms at seven. Rolled set loops keep the state in LDS rows.
constexpr int kGroupWaves = 5;

template <typename Real, typename Sh, int G>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(kGroupWaves)))
ccr_group_kernel(PathArgs<Real> a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const CCR_CONST GlobalDev& g = *toConst(a.g);
  const int rb = xcdRemap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int ngroups = (a.nsets + G - 1) / G;
  const int grp = rb % ngroups, lb = rb / ngroups;
  const int s0 = grp * G, ns = a.nsets - s0 < G ? a.nsets - s0 : G;
  const int bd = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int p = lb * bd + tid;
  const bool active = p < g.num_paths;
  const int pe = active ? p : g.num_paths - 1;  // inactive lanes replay a valid path, contribute nothing
  const int waveLocal = lb * (bd >> 6) + (tid >> 6);
  cptr<SetDev> sets = toConst(a.sets);
  const int64_t q = g.antithetic ? (pe >> 1) : pe;
  const Real* zbase = a.stream + q * sets[s0].stream_stride;
  Real* col = reinterpret_cast<Real*>(lds_raw) + tid;
  // Per-set state rows (doubles) after the Real rows; bd is a multiple of 64, so they stay aligned.
  const int realRows = a.slots + G * (a.ring + a.price_rows);
  double* state = reinterpret_cast<double*>(lds_raw + static_cast<size_t>(realRows) * bd * sizeof(Real)) + tid;
  simulateGroup<Real, Sh, G>(toConst(a.g), sets, s0, ns, pe, zbase, toConst(a.steps), toConst(a.dates),
                             toConst(a.date_reals), toConst(a.bonds), toConst(a.fixed), toConst(a.fixstore),
                             toConst(a.poly), a.bermv, a.ring, a.price_rows, a.slots, col, state, bd,
                             [&](int k, int rr, const double* st9, double vp) {
    const CCR_CONST SetDev& S = sets[s0 + k];
    if (active) a.rows[S.rows_off + static_cast<int64_t>(rr) * g.num_paths + p] = static_cast<Real>(vp);
    double* part = a.partials + S.partial_off + static_cast<int64_t>(waveLocal) * a.max_rep * kStats;
    storeStats<Sh>(st9, active, part + rr * kStats);
  });
}

// Nested conditional initial margin (src/gpu/fused_kernels.cpp:445-479), one outer path per lane:
// the path's market and fixings as in ccr_path_kernel (replayed normals), and on every reporting
// date t with a target t + mpor, inner paths from the path's state over (t, t + mpor] --
// innerStep on the inner generator's normals, the fixings of the window's dates into the path's
// fixing rows (those rows saved before and restored after every inner path) -- valued at the
// target with the same valuation records. IM(t) is the interpolated order statistic of
// |V_inner(t + mpor) - V(t)| at the plan's level, read from the 16 largest samples (kept in
// registers, sorted by insertion). Writes IM per (reporting date, path) into rows laid out as the
// exposure rows and, per wavefront, the sums of D IM and IM into partials laid out as the
// statistics' (slots 0 and 1).
template <typename Real, typename Sh, bool kOptions>
__global__ void __launch_bounds__(kMaxBlock) ccr_nested_im_kernel(PathArgs<Real> a, NestedArgs n) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const CCR_CONST GlobalDev& g = *toConst(a.g);
  const int b = xcdRemap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int set = toConst(a.block_set)[b];
  const CCR_CONST SetDev& S = toConst(a.sets)[set];
  const int bd = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int lb = b - S.block_begin;
  const int p = lb * bd + tid;
  const bool active = p < g.num_paths;
  const int pe = active ? p : g.num_paths - 1;  // inactive lanes replay a valid path, contribute nothing
  const int waveLocal = lb * (bd >> 6) + (tid >> 6);
  double* part = n.partials + S.partial_off + static_cast<int64_t>(waveLocal) * a.max_rep * kStats;
  Real* rows = static_cast<Real*>(n.rows_raw) + S.rows_off;
  Real* xs = reinterpret_cast<Real*>(lds_raw) + tid;
  const int stride = bd;
  const int64_t q = g.antithetic ? (pe >> 1) : pe;
  const Real* zbase = a.stream + q * S.stream_stride;
  cptr<Real> steps = toConst(a.steps) + static_cast<int64_t>(S.date_off) * kStepStride;
  cptr<DateRec> dates = toConst(a.dates) + S.date_off;
  cptr<Real> dreal = toConst(a.date_reals) + static_cast<int64_t>(S.date_off) * kDateRealStride;
  cptr<Real> bonds = toConst(a.bonds) + 2 * static_cast<int64_t>(S.bond_off);
  cptr<Real> fixed = toConst(a.fixed) + 4 * static_cast<int64_t>(S.fixed_off);
  cptr<int32_t> fstore = toConst(a.fixstore) + 2 * S.fixstore_off;
  cptr<Real> poly = toConst(a.poly) + S.poly_off;
  cptr<Real> opts = kOptions ? toConst(a.opts) + kOptStride * static_cast<int64_t>(S.opt_off) : nullptr;
  cptr<int32_t> target = toConst(n.target) + S.date_off;
  const Real sgn = (g.antithetic && (pe & 1)) ? Real(-1) : Real(1);
  const uint64_t pathKey = nestedPathKey(n.seed, set, pe);
  const double rank = n.level * static_cast<double>(n.inner - 1);
  const int kq = static_cast<int>(floor(rank));
  const int hiIdx = n.inner - 1 - kq < kNestedTop - 1 ? n.inner - 1 - kq : kNestedTop - 1;
  Market<Real, Sh> mk;
  marketInit<Real, Sh, false>(mk, g, zbase, S.num_dates);
  for (int j = 0; j < S.num_dates; ++j) {
    if (j > 0) marketStep<Real, Sh, false>(mk, g, j, S.num_dates, steps, zbase, sgn);
    const CCR_CONST DateRec& rec = dates[j];
    storeFixings(mk, rec.fixstore_beg, rec.fixstore_end, fstore, xs, stride);
    if (!(rec.flags & kReporting)) continue;
    const int tgt = target[j];
    double im = 0.0;
    if (tgt >= 0) {
      const Real V = valueRecord<Real, Sh, kOptions>(mk, g, rec, dreal + static_cast<int64_t>(j) * kDateRealStride, poly,
                                                     bonds, fixed, opts, xs, stride);
      // The fixing rows the window's dates write: saved, restored after every inner path.
      const int wb = dates[j + 1].fixstore_beg, nsave = dates[tgt].fixstore_end - wb;
      Real saved[kNestedMaxSave];
#pragma unroll
      for (int k = 0; k < kNestedMaxSave; ++k)
        if (k < nsave) saved[k] = xs[fstore[2 * (wb + k) + 1] * stride];
      double top[kNestedTop];
#pragma unroll
      for (int k = 0; k < kNestedTop; ++k) top[k] = 0.0;
      for (int i = 0; i < n.inner; ++i) {
        Market<Real, Sh> mi = mk;
        const uint64_t key = nestedPairKey(pathKey, j, i >> 1);
        const double isg = (i & 1) ? -1.0 : 1.0;
        for (int jj = j + 1; jj <= tgt; ++jj) {
          innerStep<Real, Sh>(mi, g, jj, steps, key, isg);
          const CCR_CONST DateRec& rj = dates[jj];
          storeFixings(mi, rj.fixstore_beg, rj.fixstore_end, fstore, xs, stride);
        }
        const Real vi = valueRecord<Real, Sh, kOptions>(mi, g, dates[tgt], dreal + static_cast<int64_t>(tgt) * kDateRealStride,
                                                        poly, bonds, fixed, opts, xs, stride);
#pragma unroll
        for (int k = 0; k < kNestedMaxSave; ++k)
          if (k < nsave) xs[fstore[2 * (wb + k) + 1] * stride] = saved[k];
        double cur = fabs(static_cast<double>(vi) - static_cast<double>(V));
#pragma unroll
        for (int k = 0; k < kNestedTop; ++k)
          if (cur > top[k]) {
            const double t = top[k];
            top[k] = cur;
            cur = t;
          }
      }
      double ok = 0.0, ok1 = 0.0;
#pragma unroll
      for (int k = 0; k < kNestedTop; ++k) {
        if (k == hiIdx) ok = top[k];
        if (k == hiIdx - 1) ok1 = top[k];
      }
      if (hiIdx == 0) ok1 = ok;
      im = ok + (rank - static_cast<double>(kq)) * (ok1 - ok);
    }
    const int rr = rec.rep;
    if (active) rows[static_cast<int64_t>(rr) * g.num_paths + p] = static_cast<Real>(im);
    const double d = static_cast<double>(mk.D);
    const double sDim = waveSum(active ? d * im : 0.0), sIm = waveSum(active ? im : 0.0);
    const int lane = tid & 63;
    if (lane < 2) part[rr * kStats + lane] = lane == 0 ? sDim : sIm;
  }
}

template <typename Real>
hipError_t launchNested(const PathArgs<Real>& args, const NestedArgs& n, int blocks, int block, hipStream_t stream) {
  const size_t lds = static_cast<size_t>(args.slots > 0 ? args.slots : 1) * block * sizeof(Real);
  return withShape(args.nrates, args.nassets, args.ncredits, [&](auto shape) -> hipError_t {
    using Sh = decltype(shape);
    if constexpr (Sh::NA > 0) {
      if (args.has_options) {
        hipLaunchKernelGGL((ccr_nested_im_kernel<Real, Sh, true>), dim3(blocks), dim3(block), lds, stream, args, n);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((ccr_nested_im_kernel<Real, Sh, false>), dim3(blocks), dim3(block), lds, stream, args, n);
    return hipGetLastError();
  });
}
template hipError_t launchNested<double>(const PathArgs<double>&, const NestedArgs&, int, int, hipStream_t);
template hipError_t launchNested<float>(const PathArgs<float>&, const NestedArgs&, int, int, hipStream_t);

bool groupSupported(int group) {
  for (int gsz : kGroupSizes)
    if (gsz == group) return true;
  return false;
}

// States pass for American Monte Carlo: the domestic state and the deflator of every date of the
// sets that hold Bermudan swaptions (simulateStates: the domestic rate only).
// Book plans (states_set >= 0) run the blocks of one set: every set with Bermudans shares its
// grid, its draws and so its states.
template <typename Real>
__global__ void __launch_bounds__(kMaxBlock) ccr_states_kernel(PathArgs<Real> a) {
  int b = xcdRemap(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  cptr<SetDev> sp;
  if (a.states_set >= 0) {
    sp = toConst(a.sets) + a.states_set;
    b += sp->block_begin;
  } else {
    sp = toConst(a.sets) + toConst(a.block_set)[b];
  }
  if (sp->state_off < 0) return;
  const int bd = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
  const int p = (b - sp->block_begin) * bd + tid;
  const int paths = toConst(a.g)->num_paths;
  simulateStates<Real>(a, sp, p < paths ? p : paths - 1, p < paths);  // lanes past the last path replay it, write nothing
}

// Sum of the per-wave partials of every (set, reporting date, statistic) in a fixed order that
// does not depend on the path kernel's workgroup size: a workgroup takes 64 consecutive (date,
// statistic) items of one set (coalesced loads), its 4 wavefronts sum the waves w = g mod 4 in
// order, then the 4 partial sums are added in order g = 0..3. Waves past the last path hold
// zeros, which leave every sum unchanged.
constexpr int kReduceGroups = 4;
__global__ void __launch_bounds__(64 * kReduceGroups) ccr_reduce_kernel(const SetDev* sets, const double* partials,
                                                                        int nsets, int max_rep, int waves_per_set,
                                                                        double inv_paths, double* out) {
  __shared__ double lds[kReduceGroups][64];
  const int per = max_rep * kStats;
  const int blocksPerSet = (per + 63) / 64;
  const int s = static_cast<int>(blockIdx.x) / blocksPerSet;
  const int rk = (static_cast<int>(blockIdx.x) % blocksPerSet) * 64 + (static_cast<int>(threadIdx.x) & 63);
  const int g = static_cast<int>(threadIdx.x) >> 6;
  double sum = 0.0;
  if (s < nsets && rk < per) {
    const double* p = partials + sets[s].partial_off + rk;
    for (int w = g; w < waves_per_set; w += kReduceGroups) sum += p[static_cast<int64_t>(w) * per];
  }
  lds[g][threadIdx.x & 63] = sum;
  __syncthreads();
  if (g == 0 && s < nsets && rk < per) {
    double total = lds[0][threadIdx.x];
    for (int k = 1; k < kReduceGroups; ++k) total += lds[k][threadIdx.x];
    out[s * per + rk] = total * inv_paths;
  }
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

// Threads per workgroup of the select: one per eight keys of the buffer (512 for 4,096 keys,
// 1,024 for 8,192), so the buffer's sorts take four compare-exchanges per thread and pass, and the
// row pass keeps that many loads in flight while the buffer limits workgroups per CU.
template <int kPfeKeys>
constexpr int pfeThreads() {
  return kPfeKeys / 8;
}
constexpr int kPfeMaxWaves = 16;
constexpr int kPfeSample = 2048; // smallest sample of a row longer than the key buffer
constexpr int kPfeBins = 4096;   // 12-bit digits of the radix fallback (histogram aliases the keys)
constexpr int kPfeUnroll = 8;    // row elements in flight per thread

// Ascending bitonic sort of keys[0, n2), n2 a power of two, by the whole workgroup: each step
// visits the n2 / 2 compare-exchange pairs once.
template <int kThreads, typename K>
__device__ inline void bitonicSort(K* keys, int n2) {
  const int tid = static_cast<int>(threadIdx.x);
  for (int k = 2; k <= n2; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int p = tid; p < (n2 >> 1); p += kThreads) {
        const int i = ((p & ~(jj - 1)) << 1) | (p & (jj - 1));
        const int ixj = i + jj;
        const bool asc = (i & k) == 0;
        const K a = keys[i], b = keys[ixj];
        if ((a > b) == asc) {
          keys[i] = b;
          keys[ixj] = a;
        }
      }
      __syncthreads();
    }
}

__device__ inline unsigned int waveInclusiveScan(unsigned int v) {
  const int lane = static_cast<int>(threadIdx.x) & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned int t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ inline unsigned int waveTotal(unsigned int v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Appends the keys of the lanes with `take` to keys[*count ...) (at most kCap are stored; *count
// keeps counting past that). One LDS atomic per wavefront. Call with the whole wave active.
template <int kCap, typename K>
__device__ inline void appendKeys(bool take, K k, K* keys, unsigned int* count) {
  const unsigned long long m = __ballot(take);
  if (m == 0) return;
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int leader = __builtin_ctzll(m);
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(count, static_cast<unsigned int>(__popcll(m)));
  base = __shfl(base, leader);
  if (take) {
    const unsigned int pos = base + static_cast<unsigned int>(__popcll(m & ((1ull << lane) - 1ull)));
    if (pos < static_cast<unsigned int>(kCap)) keys[pos] = k;
  }
}

// appendKeys for N keys per lane at once (those with take[u], in u order): one wavefront scan of
// the lanes' counts and one LDS atomic per wavefront instead of one per key. Call with the whole
// wave active.
template <int kCap, int N, typename K>
__device__ inline void appendKeysN(const bool (&take)[N], const K (&k)[N], K* keys, unsigned int* count) {
  unsigned int n = 0;
#pragma unroll
  for (int u = 0; u < N; ++u) n += take[u] ? 1u : 0u;
  if (__ballot(n != 0) == 0) return;
  const unsigned int incl = waveInclusiveScan(n);
  const unsigned int total = __shfl(incl, 63);
  unsigned int base = 0;
  if ((threadIdx.x & 63) == 0) base = atomicAdd(count, total);
  unsigned int pos = __shfl(base, 0) + incl - n;
#pragma unroll
  for (int u = 0; u < N; ++u)
    if (take[u]) {
      if (pos < static_cast<unsigned int>(kCap)) keys[pos] = k[u];
      ++pos;
    }
}

// Finds the bin of hist[0, kPfeBins) holding rank `rank` (0-based, below the histogram's total):
// each thread owns kPfeBins / kThreads consecutive bins, a workgroup prefix sum of the per-thread counts names the
// owner, which scans its bins. Leaves the bin and the rank within it in *bin / *binRank; ends with
// a barrier.
template <int kThreads>
__device__ inline void pfeFindBin(const unsigned int* hist, long long rank, unsigned int* waveTot, int* bin,
                                  long long* binRank) {
  constexpr int per = kPfeBins / kThreads;
  const int tid = static_cast<int>(threadIdx.x), lane = tid & 63, w = tid >> 6;
  unsigned int c[per], tot = 0;
#pragma unroll
  for (int k = 0; k < per; ++k) {
    c[k] = hist[tid * per + k];
    tot += c[k];
  }
  const unsigned int incl = waveInclusiveScan(tot);
  if (lane == 63) waveTot[w] = incl;
  __syncthreads();
  long long excl = static_cast<long long>(incl) - tot;
  for (int k = 0; k < w; ++k) excl += waveTot[k];
  if (rank >= excl && rank < excl + tot) {
    long long r = rank - excl;
    int b = 0;
    for (; b < per - 1; ++b) {
      if (r < static_cast<long long>(c[b])) break;
      r -= c[b];
    }
    *bin = tid * per + b;
    *binRank = r;
  }
  __syncthreads();
}

// One workgroup per (set, reporting date) row: exact order statistics lo = floor(h), hi = lo + 1
// of the row's keys (h = q (P - 1)), then the interpolation of ccr::quantile.
//  1. A row no longer than the key buffer (4096 keys, 8192 past 100,000 paths) is sorted in LDS
//     and done. Otherwise its first 2048 keys (more when 2048 would leave too many keys between
//     the splitters; paths are exchangeable) are sorted, and two sample keys a <= b, 4.5 binomial
//     standard deviations (+2) either side of the sample's q-quantile, bracket the row's
//     statistics with high probability.
//  2. One pass over the row counts keys < a, == a and == b, keeps the smallest key > b and
//     gathers the keys strictly between a and b into LDS (wave-aggregated slots).
//  3. If lo falls in [a, b] and the gathered keys fit, they are sorted and the statistics read off
//     the virtual sorted list (a x #a, the gathered keys, b x #b, then the smallest key > b). Ties
//     at the splitters (all-equal rows, rows of zeros) cost nothing.
// Otherwise (a sample miss, or too many keys between the splitters: large P with a central
// quantile) an MSB radix select over 12-bit digits narrows the prefix until its keys fit in LDS.
template <typename Real, int kPfeKeys>
// Eight row elements in flight per lane, held to 64 VGPRs (eight waves per SIMD) so that two
// workgroups of the 8,192-key buffer share a CU.
__global__ void __launch_bounds__(pfeThreads<kPfeKeys>()) __attribute__((amdgpu_waves_per_eu(8))) ccr_pfe_kernel(const SetDev* sets, const Real* rows, int max_rep,
                                                              int num_paths, double level, double* out,
                                                              int which) {
  using K = typename KeyOf<Real>::T;
  constexpr int kBits = KeyOf<Real>::kBits;
  __shared__ K keys[kPfeKeys];
  unsigned int* hist = reinterpret_cast<unsigned int*>(keys);
  constexpr int kPfeThreads = pfeThreads<kPfeKeys>();
  static_assert(kPfeThreads / 64 <= kPfeMaxWaves && kPfeBins % kPfeThreads == 0, "select workgroup shape");
  __shared__ unsigned int sBelow, sEqA, sEqB, sOpen, sWave[kPfeMaxWaves];
  __shared__ K sMinAbove;
  __shared__ long long sRank;
  __shared__ int sBin;
  const int s = static_cast<int>(blockIdx.x) / max_rep, rr = static_cast<int>(blockIdx.x) % max_rep;
  const SetDev& S = sets[s];
  // which 0: the exposure rows (one per reporting date); 1: the initial-margin proxy's
  // |V_j - V_j-1| rows (one per simulation date after the first).
  if (rr >= (which ? S.num_dates - 1 : S.num_rep)) return;
  const Real* row = rows + (which ? S.im_off : S.rows_off) + static_cast<int64_t>(rr) * num_paths;
  const int tid = static_cast<int>(threadIdx.x), lane = tid & 63;
  const double h = level * static_cast<double>(num_paths - 1);
  const long long lo = static_cast<long long>(floor(h));
  const long long hi = lo + 1 < num_paths ? lo + 1 : num_paths - 1;
  auto finish = [&](K kLo, K kHi) {
    const double vlo = static_cast<double>(valueOf<Real>(kLo)), vhi = static_cast<double>(valueOf<Real>(kHi));
    out[s * max_rep + rr] = lerpUnfused(vlo, h - static_cast<double>(lo), vhi);
  };

  // 1. Sample.
  // The smallest sample (2048, 4096, ... keys) whose splitters bracket at most 60% of the buffer.
  auto bracketed = [&](double n) { return 2.0 * (4.5 * sqrt(n * level * (1.0 - level)) + 2.0) / n * num_paths; };
  int ns = num_paths;
  if (num_paths > kPfeKeys) {
    ns = kPfeSample;
    while (ns < kPfeKeys && bracketed(ns) > 0.6 * kPfeKeys) ns <<= 1;
  }
  int n2 = 1;
  while (n2 < ns) n2 <<= 1;
  for (int i = tid; i < n2; i += kPfeThreads) keys[i] = i < ns ? keyOf<Real>(row[i]) : ~K(0);
  __syncthreads();
  bitonicSort<kPfeThreads>(keys, n2);
  if (ns == num_paths) {
    if (tid == 0) finish(keys[lo], keys[hi]);
    return;
  }
  const long long ls = static_cast<long long>(floor(level * static_cast<double>(ns - 1)));
  const long long margin = static_cast<long long>(ceil(4.5 * sqrt(ns * level * (1.0 - level)))) + 2;
  const K a = ls - margin >= 0 ? keys[ls - margin] : K(0);
  const K b = ls + 1 + margin < ns ? keys[ls + 1 + margin] : ~K(0);  // ~0: no upper splitter
  const bool twoEnds = b != a;
  if (tid == 0) {
    sBelow = sEqA = sEqB = sOpen = 0;
    sMinAbove = ~K(0);
  }
  __syncthreads();  // the splitters are read before the key buffer is reused

  // 2. One pass over the row.
  {
    unsigned int below = 0, eqA = 0, eqB = 0;
    K minAbove = ~K(0);
    // Full spans of kPfeThreads x kPfeUnroll keys need no bounds tests; the last, partial span
    // marks the keys past the row ~0 and counts them nowhere.
    constexpr int kSpan = kPfeThreads * kPfeUnroll;
    const int fullEnd = num_paths / kSpan * kSpan;
    auto visit = [&](int base, auto tailTag) {
      constexpr bool kTail = decltype(tailTag)::value;
      K kk[kPfeUnroll];
      bool take[kPfeUnroll];
#pragma unroll
      for (int u = 0; u < kPfeUnroll; ++u) {
        const int i = base + u * kPfeThreads + tid;
        kk[u] = (!kTail || i < num_paths) ? keyOf<Real>(row[i]) : ~K(0);
      }
#pragma unroll
      for (int u = 0; u < kPfeUnroll; ++u) {
        const K k = kk[u];
        const bool valid = !kTail || k != ~K(0);
        below += (valid && k < a) ? 1u : 0u;
        eqA += (valid && k == a) ? 1u : 0u;
        eqB += (valid && twoEnds && k == b) ? 1u : 0u;
        if (valid && k > b && k < minAbove) minAbove = k;
        take[u] = valid && k > a && k < b;
      }
      appendKeysN<kPfeKeys>(take, kk, keys, &sOpen);
    };
    for (int base = 0; base < fullEnd; base += kSpan) visit(base, std::false_type{});
    if (fullEnd < num_paths) visit(fullEnd, std::true_type{});
    below = waveTotal(below);
    eqA = waveTotal(eqA);
    eqB = waveTotal(eqB);
    if (lane == 0) {
      atomicAdd(&sBelow, below);
      atomicAdd(&sEqA, eqA);
      atomicAdd(&sEqB, eqB);
    }
    atomicMin(&sMinAbove, minAbove);
  }
  __syncthreads();
  {
    const long long nBelow = sBelow, nA = sEqA, nB = sEqB;
    const unsigned int nOpen = sOpen;
    if (nOpen <= static_cast<unsigned int>(kPfeKeys) && lo >= nBelow && lo < nBelow + nA + nOpen + nB) {
      // 3. Sort the gathered keys and read the statistics off the virtual list.
      int m2 = 1;
      while (m2 < static_cast<int>(nOpen)) m2 <<= 1;
      for (int i = static_cast<int>(nOpen) + tid; i < m2; i += kPfeThreads) keys[i] = ~K(0);
      __syncthreads();
      bitonicSort<kPfeThreads>(keys, m2);
      if (tid == 0) {
        auto at = [&](long long r, K* v) {
          long long i = r - nBelow;
          if (i < nA) {
            *v = a;
            return true;
          }
          i -= nA;
          if (i < static_cast<long long>(nOpen)) {
            *v = keys[i];
            return true;
          }
          i -= nOpen;
          if (i < nB) {
            *v = b;
            return true;
          }
          return false;
        };
        K kLo = 0, kHi = 0;
        at(lo, &kLo);
        if (hi == lo) kHi = kLo;
        else if (!at(hi, &kHi)) kHi = sMinAbove;  // hi is the first key past b
        finish(kLo, kHi);
      }
      return;
    }
  }

  // Fallback: MSB radix select over 12-bit digits (the last digit takes the remaining bits).
  K prefix = 0, mask = 0;
  long long rank = lo;
  for (int shift = kBits; shift > 0;) {
    const int width = shift >= 12 ? 12 : shift;
    shift -= width;
    const K digit = (K(1) << width) - 1;
    __syncthreads();
    for (int i = tid; i < kPfeBins; i += kPfeThreads) hist[i] = 0;
    __syncthreads();
    for (int i = tid; i < num_paths; i += kPfeThreads) {
      const K k = keyOf<Real>(row[i]);
      if ((k & mask) == prefix) atomicAdd(&hist[static_cast<int>((k >> shift) & digit)], 1u);
    }
    __syncthreads();
    pfeFindBin<kPfeThreads>(hist, rank, sWave, &sBin, &sRank);
    const int bin = sBin;
    const unsigned int cnt = hist[bin];
    rank = sRank;
    prefix |= static_cast<K>(bin) << shift;
    mask |= digit << shift;
    if (shift > 0 && cnt > static_cast<unsigned int>(kPfeKeys)) continue;
    // The keys under the prefix (all equal once shift is 0) fit in LDS: gather them with the
    // smallest key above the prefix, sort, read off the statistics.
    __syncthreads();
    if (tid == 0) {
      sOpen = 0;
      sMinAbove = ~K(0);
    }
    __syncthreads();
    K minAbove = ~K(0);
    for (int base = 0; base < num_paths; base += kPfeThreads) {
      const int i = base + tid;
      const K k = i < num_paths ? keyOf<Real>(row[i]) : K(0);
      const bool in = i < num_paths && (k & mask) == prefix;
      if (i < num_paths && (k & mask) > prefix && k < minAbove) minAbove = k;
      appendKeys<kPfeKeys>(in, k, keys, &sOpen);
    }
    atomicMin(&sMinAbove, minAbove);
    __syncthreads();
    if (cnt > static_cast<unsigned int>(kPfeKeys)) {  // shift is 0: every key under the prefix equals it
      if (tid == 0) finish(prefix, (hi == lo || rank + 1 < static_cast<long long>(cnt)) ? prefix : sMinAbove);
      return;
    }
    int m2 = 1;
    while (m2 < static_cast<int>(cnt)) m2 <<= 1;
    for (int i = static_cast<int>(cnt) + tid; i < m2; i += kPfeThreads) keys[i] = ~K(0);
    __syncthreads();
    bitonicSort<kPfeThreads>(keys, m2);
    if (tid == 0) {
      const K kLo = keys[rank];
      const K kHi = hi == lo ? kLo : (rank + 1 < static_cast<long long>(cnt) ? keys[rank + 1] : sMinAbove);
      finish(kLo, kHi);
    }
    return;
  }
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
// Per-trade plans run the generic shape with options and per-trade values compiled in; their
// LDS column adds one row per trade.
template <typename Real, typename Sh>
hipError_t launchGroupShape(const PathArgs<Real>& args, int block, size_t lds, hipStream_t stream) {
  const int ngroups = (args.nsets + args.group - 1) / args.group;
  const dim3 grid(static_cast<unsigned>(ngroups * args.blocks_per_set));
  switch (args.group) {
    case 2: hipLaunchKernelGGL((ccr_group_kernel<Real, Sh, 2>), grid, dim3(block), lds, stream, args); break;
    case 4: hipLaunchKernelGGL((ccr_group_kernel<Real, Sh, 4>), grid, dim3(block), lds, stream, args); break;
    case 8: hipLaunchKernelGGL((ccr_group_kernel<Real, Sh, 8>), grid, dim3(block), lds, stream, args); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename Real>
hipError_t launchPath(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream) {
  if (args.group > 1) {
    const size_t lds = static_cast<size_t>(args.slots + args.group * (args.ring + args.price_rows)) * block * sizeof(Real) +
                       (!groupUnrolled(args.group) ? static_cast<size_t>(kSetStateRows * args.group) * block * sizeof(double) : 0);
    return withShape(args.nrates, args.nassets, args.ncredits, [&](auto shape) {
      return launchGroupShape<Real, decltype(shape)>(args, block, lds, stream);
    });
  }
  if (args.tstride > 0) {
    const size_t lds = static_cast<size_t>(args.ring + args.price_rows + args.slots + args.tstride / kTradeStats) *
                       block * sizeof(Real);
    hipLaunchKernelGGL((ccr_path_kernel<Real, GenericShape, true, true>), dim3(blocks), dim3(block), lds, stream, args);
    return hipGetLastError();
  }
  const size_t lds = static_cast<size_t>(args.ring + args.price_rows + args.slots) * block * sizeof(Real);
  return withShape(args.nrates, args.nassets, args.ncredits, [&](auto shape) {
    return launchShape<Real, decltype(shape)>(args, blocks, block, lds, stream);
  });
}

template <typename Real>
hipError_t pathOccupancy(int nrates, int nassets, int ncredits, bool options, int ldsRows, int block, int* wgPerCu,
                         bool trades, int group) {
  // Group kernels: ldsRows Real rows plus, where the set loop is not unrolled, kSetStateRows
  // double rows per set.
  const size_t lds = static_cast<size_t>(ldsRows) * block * sizeof(Real) +
                     (!groupUnrolled(group) ? static_cast<size_t>(kSetStateRows * group) * block * sizeof(double) : 0);
  if (group > 1)
    return withShape(nrates, nassets, ncredits, [&](auto shape) -> hipError_t {
      using Sh = decltype(shape);
      switch (group) {
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_group_kernel<Real, Sh, 2>, block, lds);
        case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_group_kernel<Real, Sh, 4>, block, lds);
        case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_group_kernel<Real, Sh, 8>, block, lds);
        default: return hipErrorInvalidValue;
      }
    });
  if (trades)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_path_kernel<Real, GenericShape, true, true>, block,
                                                        lds);
  return withShape(nrates, nassets, ncredits, [&](auto shape) -> hipError_t {
    using Sh = decltype(shape);
    if constexpr (Sh::NA > 0) {
      if (options)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_path_kernel<Real, Sh, true>, block, lds);
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(wgPerCu, ccr_path_kernel<Real, Sh, false>, block, lds);
  });
}
template hipError_t pathOccupancy<double>(int, int, int, bool, int, int, int*, bool, int);
template hipError_t pathOccupancy<float>(int, int, int, bool, int, int, int*, bool, int);

__global__ void __launch_bounds__(256) ccr_narrow_kernel(const double* in, float* out, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256)
    out[i] = static_cast<float>(in[i]);
}

hipError_t launchNarrow(const double* in, float* out, int64_t n, hipStream_t stream) {
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(ccr_narrow_kernel, dim3(static_cast<unsigned>(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                     stream, in, out, n);
  return hipGetLastError();
}

// dom[dom_off + (2 s + i) Q + q] = stream[q L + s nz + (i ? ncorr : 0)] for every set with a
// domestic-normals slice, step s < T - 1 and path block q < Q (blockIdx.z: the set's position
// among those sets).
template <typename Real>
__global__ void __launch_bounds__(256) ccr_domestic_normals_kernel(const SetDev* sets, const int32_t* which,
                                                                    const Real* stream, Real* dom, int64_t Q, int nz,
                                                                    int ncorr) {
  const SetDev& S = sets[which[blockIdx.z]];
  const int s = static_cast<int>(blockIdx.y);
  if (s >= S.num_dates - 1) return;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (q >= Q) return;
  const Real* z = stream + q * S.stream_stride + static_cast<int64_t>(s) * nz;
  Real* d = dom + S.dom_off + static_cast<int64_t>(2 * s) * Q + q;
  d[0] = z[0];
  d[Q] = z[ncorr];
}

template <typename Real>
hipError_t launchDomesticNormals(const PathArgs<Real>& args, const SetDev* hostSets, int nsets, Real* dom,
                                 hipStream_t stream) {
  // One gather per distinct slice (the sets of a book share one).
  std::vector<int32_t> which;
  std::vector<int64_t> seen;
  int maxDates = 0;
  for (int s = 0; s < nsets; ++s)
    if (hostSets[s].dom_off >= 0) {
      bool dup = false;
      for (int64_t o : seen) dup = dup || o == hostSets[s].dom_off;
      if (dup) continue;
      seen.push_back(hostSets[s].dom_off);
      which.push_back(s);
      maxDates = hostSets[s].num_dates > maxDates ? hostSets[s].num_dates : maxDates;
    }
  if (which.empty() || maxDates < 2) return hipSuccess;
  GlobalDev g;
  hipError_t e = hipMemcpy(&g, args.g, sizeof(GlobalDev), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  const int64_t Q = g.antithetic ? (static_cast<int64_t>(g.num_paths) + 1) / 2 : g.num_paths;
  int32_t* dwhich = nullptr;
  if ((e = hipMalloc(&dwhich, which.size() * sizeof(int32_t))) != hipSuccess) return e;
  e = hipMemcpy(dwhich, which.data(), which.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    for (size_t b = 0; b < which.size() && e == hipSuccess; b += 65535) {
      const unsigned nz3 = static_cast<unsigned>(which.size() - b < 65535 ? which.size() - b : 65535);
      hipLaunchKernelGGL(ccr_domestic_normals_kernel<Real>, dim3(static_cast<unsigned>((Q + 255) / 256), maxDates - 1, nz3),
                         dim3(256), 0, stream, args.sets, dwhich + b, args.stream, dom, Q, g.nz, g.ncorr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
  }
  const hipError_t f = hipFree(dwhich);
  return e != hipSuccess ? e : f;
}
template hipError_t launchDomesticNormals<double>(const PathArgs<double>&, const SetDev*, int, double*, hipStream_t);
template hipError_t launchDomesticNormals<float>(const PathArgs<float>&, const SetDev*, int, float*, hipStream_t);

template <typename Real>
hipError_t launchStates(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream) {
  hipLaunchKernelGGL(ccr_states_kernel<Real>, dim3(blocks), dim3(block), 0, stream, args);
  return hipGetLastError();
}
template hipError_t launchStates<double>(const PathArgs<double>&, int, int, hipStream_t);
template hipError_t launchStates<float>(const PathArgs<float>&, int, int, hipStream_t);

hipError_t launchReduce(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                        double inv_paths, double* out, hipStream_t stream) {
  const int blocks = nsets * ((max_rep * kStats + 63) / 64);
  hipLaunchKernelGGL(ccr_reduce_kernel, dim3(blocks), dim3(64 * kReduceGroups), 0, stream, sets, partials, nsets,
                     max_rep, waves_per_set, inv_paths, out);
  return hipGetLastError();
}

template <typename Real>
hipError_t launchPfe(const SetDev* sets, const Real* rows, int nsets, int max_rep, int num_paths, double level,
                     double* out, hipStream_t stream, int which) {
  // Rows beyond 100,000 paths take an 8,192-key buffer, so the keys between the splitters of a
  // 4,096-key sample still fit.
  if (num_paths > 100000)
    hipLaunchKernelGGL((ccr_pfe_kernel<Real, 8192>), dim3(nsets * max_rep), dim3(pfeThreads<8192>()), 0, stream, sets,
                       rows, max_rep, num_paths, level, out, which);
  else
    hipLaunchKernelGGL((ccr_pfe_kernel<Real, 4096>), dim3(nsets * max_rep), dim3(pfeThreads<4096>()), 0, stream, sets,
                       rows, max_rep, num_paths, level, out, which);
  return hipGetLastError();
}

// Sum of the per-wave deflator sums of every (set, simulation date), in wave order.
__global__ void ccr_im_reduce_kernel(const SetDev* sets, const double* part, int nsets, int dates, int waves_per_set,
                                     double* out) {
  const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= nsets * dates) return;
  const int s = idx / dates, j = idx % dates;
  const double* p = part + sets[s].im_part_off + j;
  double sum = 0.0;
  for (int w = 0; w < waves_per_set; ++w) sum += p[static_cast<int64_t>(w) * dates];
  out[idx] = sum;
}

hipError_t launchImReduce(const SetDev* sets, const double* part, int nsets, int dates, int waves_per_set, double* out,
                          hipStream_t stream) {
  const int n = nsets * dates;
  hipLaunchKernelGGL(ccr_im_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, sets, part, nsets, dates,
                     waves_per_set, out);
  return hipGetLastError();
}

// Per-trade statistics: the fixed-order sum over a set's waves of every (reporting date, value)
// of its partials, divided by the paths.
__global__ void ccr_trade_reduce_kernel(const SetDev* sets, const double* part, int nsets, int max_rep, int tstride,
                                        int waves_per_set, double inv_paths, double* out) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t per = static_cast<int64_t>(max_rep) * tstride;
  if (idx >= nsets * per) return;
  const int s = static_cast<int>(idx / per);
  const int64_t e = idx - s * per;
  const double* p = part + sets[s].tpart_off + e;
  double sum = 0.0;
  for (int w = 0; w < waves_per_set; ++w) sum += p[static_cast<int64_t>(w) * per];
  out[idx] = sum * inv_paths;
}

hipError_t launchTradeReduce(const SetDev* sets, const double* part, int nsets, int max_rep, int tstride,
                             int waves_per_set, int num_paths, double* out, hipStream_t stream) {
  const int64_t n = static_cast<int64_t>(nsets) * max_rep * tstride;
  hipLaunchKernelGGL(ccr_trade_reduce_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, sets,
                     part, nsets, max_rep, tstride, waves_per_set, 1.0 / static_cast<double>(num_paths), out);
  return hipGetLastError();
}

template hipError_t launchPath<double>(const PathArgs<double>&, int, int, hipStream_t);
template hipError_t launchPath<float>(const PathArgs<float>&, int, int, hipStream_t);
template hipError_t launchPfe<double>(const SetDev*, const double*, int, int, int, double, double*, hipStream_t, int);
template hipError_t launchPfe<float>(const SetDev*, const float*, int, int, int, double, double*, hipStream_t, int);

}  // namespace ccrhip
