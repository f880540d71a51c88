// launch.hpp -- kernel arguments and host-side launchers of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

// Plan tables are read through the constant address space on the device: AMDGPU treats such
// loads as invariant and issues them as scalar (s_load) loads when the address is uniform, so
// one load feeds all 64 lanes. Host code (the CPU emulator) sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define CCR_CONST __attribute__((address_space(4)))
#else
#define CCR_CONST
#endif

namespace ccrhip {

template <class T>
using cptr = const CCR_CONST T*;
template <class T>
__host__ __device__ inline cptr<T> toConst(const T* p) {
  return (cptr<T>)p;  // address-space cast (a C-style cast is the form clang accepts)
}

// The regression engine's continuation fits as the valuation reads them: with the swap-rate /
// annuity basis, per (swaption, date) record its fit record rcont[record] (rfit, kBasisStride
// reals) and the tail's annuity / float-leg terms; otherwise (basis < 0 or the state basis) the
// per-record fits of `cont`.
struct BasisView {
  int basis = -1;
  cptr<RegRec> rrecs = nullptr;
  cptr<int32_t> rcont = nullptr;
  cptr<double> rfit = nullptr;
  cptr<TermRange> saterms = nullptr;
};

template <typename Real>
struct PathArgs {
  const GlobalDev* g;
  const SetDev* sets;
  const int32_t* block_set;
  const Real* stream;
  const Real* dom;     // AMC states pass: per set [step][2][path block] the domestic driver and integral normals
  const Real* steps;
  const DateRec* dates;
  const Real* date_reals;
  const Real* bonds;
  const Real* fixed;
  const int32_t* fixstore;
  const Real* poly;    // bond-polynomial coefficients in the plan's precision
  const Real* opts;    // European option terms
  const double* bermv;      // American Monte Carlo: [continuation date][path] sum of the set's swaptions
  double* states;           // states-only pass: [set dates][path] x then D (null otherwise)
  int32_t states_set;       // states pass of a book plan: the one set it runs (-1: every set's blocks)
  Real* rows;
  double* partials;
  int32_t max_rep;
  int32_t ring;        // LDS rows of the margin-call value ring
  int32_t price_rows;  // LDS rows of the collateral price ring (0 for cash collateral)
  int32_t slots;       // LDS rows of fixing states
  int32_t nrates, nassets, ncredits;  // model shape: selects the kernel specialisation
  int32_t has_options;                // European options present: selects the kernel variant
  int32_t group;                      // book plans: netting sets per workgroup (0: one set per workgroup)
  int32_t blocks_per_set;             // path blocks of every set (group launches)
  int32_t nsets;
  int64_t state_len;                  // elements of x in `states` (D follows)
  // Initial-margin proxy (null when off): |V_j - V_j-1| rows per set from SetDev::im_off and
  // per-wave deflator sums from SetDev::im_part_off, im_dates dates per row block.
  Real* im_rows;
  double* im_part;
  int32_t im_dates;
  // Per-trade plans (tstride > 0): per-(date, trade) valuation records and reals, and per-wave
  // partials [set waves][R_max][tstride] of the kTradeStats statistics of every trade.
  int32_t tstride;
  const DateRec* tdates;
  const Real* tdreals;
  double* tpart;
};

// American Monte Carlo regressions (amc.hip). Layouts: states x then D, [set dates][path];
// ev and strategy [global exercise][path]; part [item][chunk][moments]; W [cdate][4][path];
// gpart [path chunk][date record][4] (continuation X'y partials); cont [date record][kContStride];
// bermv [cdate][path].
struct AmcArgs {
  const SetDev* sets;
  const BermDev* berms;
  const BermDate* bdates;
  const TermRange* evterms;
  const int32_t* eidx;        // exercise date index per (Bermudan, exercise)
  const int32_t* exlist;      // Bermudan of each global exercise
  const int32_t* cdates;      // (set, date) per continuation date
  const LiveRec* clive;       // swaptions with a value, per continuation date ...
  const int32_t* clive_rng;   // ... [begin, end) pairs
  const int32_t* bdate_berm;  // Bermudan of each date record
  const int32_t* bdate_cd;    // continuation date of each date record (-1: none)
  const double* poly;         // bond-polynomial coefficients of FP64 plans ...
  const float* polyf;         // ... and of FP32 plans
  const double* states;
  double* ev;
  double* strategy;
  double* part;
  double* regp;               // [Bermudan][8]: n, center, scale, degree, sum y
  double* coef;               // [Bermudan][4]
  double* csum;               // [cdate][3]: n, sum x, sum x^2
  double* gram;               // [cdate][10]
  const GemmTile* tiles;
  const GemmTile* narrow;     // narrow tiles
  const int32_t* gcols;       // strategy column per (tile, swaption), -1: none
  const int32_t* gberm;       // global Bermudan per (tile, swaption)
  const int32_t* tcd;         // continuation date per (tile, date)
  const int32_t* cmap;        // continuation date holding the moments and Gram of each one
  const int32_t* cuniq;       // the continuation dates that hold them (ncuniq)
  double* gpart;
  double* cont;
  int16_t* exidx;             // [Bermudan][path]
  double* bermv;
  int64_t paths, state_len;
  int32_t nberms, ncols, ncdates, nbdates, nchunks, ntiles, nkc;
  int32_t nnarrow;
  int32_t ncchunks;           // continuation-date partials per date: ceil(nchunks / kContGroup)
  int32_t ncuniq;
  // Regression engine with a basis choice (basis >= 0): records, exercise and continuation
  // look-ups, the tails' annuity / float-leg terms, per-record moments / Gram partials
  // [record][chunk][kBasisGram] and fits [record][kBasisStride].
  int32_t basis, gram32, nrecs, nex_recs, nbc, max_exercises;
  const RegRec* rrecs;
  const int32_t* rex;
  const int32_t* rcont;
  const TermRange* saterms;
  double* rpart;
  double* rfit;
};

template <typename Real>
hipError_t launchAmc(const AmcArgs& args, const Real* bonds, const Real* fixed, int max_exercises, hipStream_t stream);

// Surrogate features of a group of netting sets with Bermudans that share one set of recorded
// states (one set, or every such set of a book plan): rows [u][path][kSurrogateBase + count] =
// (t, x, running max x, running min x, x at each exercise date once passed, else 0) (FP32) for
// the group's rows u: date j of the states is row urow[j] (-1: no row), at time utimes[u].
constexpr int kSurrogateBase = 4;
constexpr int kSurrogateMaxExercises = 28;  // BF16 network inputs are at most 32
struct SurrogateExercises {
  int32_t count;                           // distinct exercise dates used (the earliest ones)
  int32_t date[kSurrogateMaxExercises];    // their indices on the simulation grid of the states
};
hipError_t launchSurrogateFeatures(const double* states, int64_t state_off, int64_t paths, const int32_t* urow,
                                   const double* utimes, int num_dates, const SurrogateExercises& ex, float* feats,
                                   hipStream_t stream);
// out[0, n) = 0.
hipError_t launchZeroRows(double* out, int64_t n, hipStream_t stream);

// Nested conditional initial margin (ccr_nested_im_kernel): per set date [date_off + j] the date
// of t_j + mpor (-1: none), inner paths, the quantile level, the simulation seed; IM rows laid out
// as the exposure rows and per-wave partials laid out as the statistics' (slots 0: D IM, 1: IM).
constexpr int kNestedTop = 16;  // largest inner samples kept per (path, date)
struct NestedArgs {
  const int32_t* target;
  int32_t inner;
  double level;
  uint64_t seed;
  void* rows_raw;  // IM rows in the plan's precision
  double* partials;
};
template <typename Real>
hipError_t launchNested(const PathArgs<Real>& args, const NestedArgs& n, int blocks, int block, hipStream_t stream);

template <typename Real>
hipError_t launchPath(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream);
// out[i] = (float)in[i], i < n (round to nearest even, as a host static_cast).
hipError_t launchNarrow(const double* in, float* out, int64_t n, hipStream_t stream);
// Gathers the two normals per step the AMC states pass reads (the first correlated driver and the
// domestic integral normal) of every set with Bermudans from the stream into args.dom.
template <typename Real>
hipError_t launchDomesticNormals(const PathArgs<Real>& args, const SetDev* hostSets, int nsets, Real* dom,
                                 hipStream_t stream);
// Workgroups of the path kernel resident per CU (registers, LDS, workgroup slots) for a model
// shape, the option variant, ldsRows LDS rows per thread and a workgroup size (group > 1: the
// book kernel with that many netting sets per workgroup).
template <typename Real>
hipError_t pathOccupancy(int nrates, int nassets, int ncredits, bool options, int ldsRows, int block, int* wgPerCu,
                         bool trades = false, int group = 0);
// Netting sets per workgroup the book kernel is compiled for (1 = the per-set kernel).
constexpr int kGroupSizes[] = {2, 4, 8};
bool groupSupported(int group);
// The states-only pass of the path kernel (AMC): writes x and D of every date into args.states.
template <typename Real>
hipError_t launchStates(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream);
hipError_t launchReduce(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                        double inv_paths, double* out, hipStream_t stream);
template <typename Real>
hipError_t launchPfe(const SetDev* sets, const Real* rows, int nsets, int max_rep, int num_paths, double level,
                     double* out, hipStream_t stream, int which = 0);
// Per (set, simulation date) deflator sums of the initial-margin proxy from its per-wave sums.
hipError_t launchTradeReduce(const SetDev* sets, const double* part, int nsets, int max_rep, int tstride,
                             int waves_per_set, int num_paths, double* out, hipStream_t stream);
hipError_t launchImReduce(const SetDev* sets, const double* part, int nsets, int dates, int waves_per_set, double* out,
                          hipStream_t stream);

}  // namespace ccrhip
