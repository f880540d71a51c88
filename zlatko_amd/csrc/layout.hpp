// layout.hpp -- the device-side layout of a compiled exposure plan, shared by the host plan
// builder and the gfx950 kernels.
//
// HBM layout (one plan, S netting sets, P paths, nz normals per step):
//   stream   Real[ceil(P/2) * Lmax]    the reference normal stream in draw order: path block q
//                                      (paths 2q, 2q+1 when antithetic) owns (T_s - 1) * nz
//                                      consecutive normals of netting set s, step-major
//   steps    Real[sum_s (T_s) * kStepStride]  per-date model coefficients (row j: step t_{j-1} -> t_j)
//   dates    DateRec[sum_s T_s] + Real[sum_s T_s * kDateRealStride]  per-date valuation records
//   bonds    Real2[...]   (c A(t,T), B(t,T)) merged per (rate, maturity) and date
//   fixed    Real4[...]   (c R A(t,te), B(t,te), B(tk,ts) - B(tk,te), slot) per fixed floating coupon
//   rows     Real[S][R_max][P]        max(V - C, 0) on the reporting dates (PFE input)
//   partials double[S][waves][R_max][kStats]  per-wavefront sums, reduced in a fixed order
#pragma once

#include <cstdint>

#include "../../include/ccr_hip.h"

namespace ccrhip {

constexpr int kMaxRates = 4;      // domestic + up to 3 foreign Hull-White rates
constexpr int kMaxAssets = 4;     // lognormal FX / equity factors
constexpr int kMaxCredits = 2;    // CIR++ intensities (counterparty, own)
constexpr int kMaxCorr = kMaxRates + kMaxAssets + kMaxCredits;
// Fixing states a nested initial-margin window (t, t + mpor] may overwrite (the nested kernel
// saves and restores them around every inner path), at most.
constexpr int kNestedMaxSave = 16;

// Book kernels whose loop over a workgroup's netting sets is unrolled, the per-set collateral
// and discounted-exposure state in registers instead of LDS rows: four sets per workgroup (two
// sets fit five waves per SIMD with the state in LDS; eight would hold too much code).
constexpr bool groupUnrolled(int group) { return group == 4; }
// Double rows per set holding that state when the set loop is rolled: units, previous V - C.
constexpr int kSetStateRows = 2;
constexpr int kMaxSub = 8;        // CIR++ sub-steps per grid step
constexpr int kStats = 9;         // E[V-C], EE, ENE, EE*, ENE*, E[D], E[Q], CVA incr., DVA incr.
constexpr int kMaxBlock = 256;    // launch bound of the path kernel (threads per workgroup)
// Per-trade statistics (allocation): at most kMaxTradeGroups trades per netting set, kTradeStats
// per trade and reporting date -- E[D V_i 1{V>0}], E[V_i 1{V>0}], E[D (V+ - (V - V_i)+)],
// E[D V_i+], E[V_i+] with V the netting set's value before collateral.
constexpr int kMaxTradeGroups = 16;
constexpr int kTradeStats = 5;

// Step record (Real[kStepStride]) for the step t_{j-1} -> t_j (row 0 unused):
//   rate r:   8 r + {e, sdX, c1, c2, (1-e)/a, drift (1-e)/a, drift (dt-(1-e)/a)/a, int phi}
//   asset a:  32 + 3 a + {int carry, vol sqrt(dt), vol^2 dt / 2}
//   credit c: 44 + 20 c + {sqrt(dt), h, shift(t_j), bridge mean k (8), bridge sd k (8), pad}
constexpr int kStepRate = 0;
constexpr int kStepAsset = 32;
constexpr int kStepCredit = 44;
constexpr int kStepCreditStride = 20;
constexpr int kStepStride = 84;

// Per-date valuation record.
struct DateRec {
  int32_t flags;        // kValued | kReporting | kStoreValue
  int32_t rep;          // reporting index (or -1)
  int32_t call_pos;     // CSA: ring position of the margin-call date's value
  int32_t store_pos;    // CSA: ring position this date's value is stored at (or -1)
  int32_t bond_beg[kMaxRates + 1];   // bond terms of rate r: [bond_beg[r], bond_beg[r + 1])
  int32_t fix_beg[kMaxRates + 1];    // fixed-coupon terms of rate r
  int32_t fixstore_beg, fixstore_end;  // (rate, slot) pairs whose state is recorded at this date
  // Bond polynomials of all rates, interleaved: poly_rows rows (a multiple of kPolyBatch) of nrates
  // coefficients from poly_beg; row k holds power poly_rows - 1 - k of every rate (zero-padded at
  // the top). 0 rows: none (the exponential terms are evaluated).
  int32_t poly_beg, poly_rows;
  int32_t pad_poly[2 * kMaxRates - 2];
  int32_t opt_beg, opt_end;          // European options alive at this date: [opt_beg, opt_end)
  int32_t vidx;                      // index of this date among the set's valued dates (-1: not valued)
  int32_t pad;
};
static_assert(sizeof(DateRec) == 112, "DateRec must stay 112 bytes");

// European option term (Real[kOptStride]) on one date: asset, +1 call / -1 put, notional,
// strike, A(t,T), B(t,T) of the domestic zero bond to expiry, carry Q(t,T), vol sqrt(T - t).
constexpr int kOptStride = 8;
constexpr int32_t kValued = 1, kReporting = 2, kStoreValue = 4;

// Per-date reals: c0[rate] (constant bond terms), lin[asset] (deterministic-carry forwards),
// collateral bond factor (A, B), |x_r| bound of each rate's bond polynomial.
constexpr int kDateC0 = 0;
constexpr int kDateLin = kMaxRates;
constexpr int kDateBond = kMaxRates + kMaxAssets;
constexpr int kDateXmax = kDateBond + 2;
constexpr int kDateRealStride = 16;

// Bond polynomial. On a valued date the merged bond terms of rate r are an entire function of
// the state, sum_T c_T A_T e^{-B_T x} = sum_k C_k x^k with C_k = sum_T c_T A_T (-B_T)^k / k!,
// whose coefficients do not depend on the path. The host truncates the series where its tail is
// below kPolyTolerance of sum_T |c_T A_T| for |x| <= xmax (a bound of 9 standard deviations plus
// the drift); a wavefront with any lane beyond xmax evaluates the exponential terms instead.
constexpr double kPolyTolerance = 1e-18;
// FP32 plans: a tail below 1e-9 of sum_T |c_T A_T| is under the rounding of their FP32 coefficients.
constexpr double kPolyToleranceFp32 = 1e-9;
constexpr int kPolyBatch = 8;  // coefficient rows per batch of scalar loads in the path kernel
constexpr double kPolySigmas = 9.0;
constexpr int kPolyMaxDegree = 48;

// ---------------------------------------------------------------- American Monte Carlo
// Bermudan swaptions are valued by Longstaff-Schwartz regressions on the domestic state
// (src/instruments/bermudan_swaption.cpp:49-146): a states-only pass of the path kernel records
// x and D per date, batched regression kernels fit the exercise policy backwards over the exercise
// dates and the continuation value on every valued date, a valuation kernel sums each set's
// swaptions per (valued date, path) -- the underlying swap once exercised (physical; 0 for cash),
// else max(continuation, 0), times the position -- and the path kernel adds that sum to the
// netting-set value.
constexpr int kMaxExercises = 64;
constexpr int kAmcChunk = 256;   // paths per workgroup of the regression reductions
constexpr int kContGroup = 16;   // path chunks per workgroup of the continuation-date reductions
constexpr int kRegMom1 = 4;      // masked moments: n, sum x, sum x^2, sum y
constexpr int kRegMom2 = 14;     // Gram lower triangle (10) + X'y (4)
constexpr int kContStride = 8;   // per (Bermudan, date): center, scale, c0..c3, degree, pad
// Continuation right-hand sides X'y for every (Bermudan, valued date): a gathered GEMM per
// exercise epoch (run of valued dates on which every swaption of the set has the same next
// exercise), tiles of kGemmDates dates x kGemmCols swaptions, split over paths in kGemmK chunks.
constexpr int kGemmDates = 8;
constexpr int kGemmCols = 64;
constexpr int kGemmK = 4096;
// Sets with at most kNarrowCols Bermudans use narrow tiles instead: up to kNarrowDates dates of
// one epoch x their (<= 8) swaptions, one lane per (date, swaption).
constexpr int kNarrowDates = 16;
constexpr int kNarrowCols = 8;

// Value terms of one instrument on one date: bond polynomial (or its exponential terms) plus
// fixed coupons; indices relative to the set's poly / bonds / fixed tables. The fixed coupons of
// a TermRange carry the fixing date index (states row) in their fourth element.
struct TermRange {
  int32_t poly_beg, poly_len;
  int32_t bond_beg, bond_end;
  int32_t fix_beg, fix_end;
  double c0, xmax;
};

// Regression engine with a choice of basis (config 3's swap-rate / annuity regressions): one
// record per regression -- the exercise decision at exercise k of a swaption (in-the-money paths,
// target realised / D) or its continuation on a valued date (all paths, target strategy[k] / D,
// k the next exercise) -- regressed on
//   kBasisState:            1, z, z^2, z^3 of the standardised domestic state (the reference's
//                           PolynomialRegression, bermudan_swaption.cpp:77-93);
//   kBasisSwapRateAnnuity:  the 10 monomials of total degree <= 3 in the standardised swap rate S
//                           and annuity A of the swaption's tail from exercise k:
//                           1, s, a, s^2, s a, a^2, s^3, s^2 a, s a^2, a^3.
// The Gram of [phi | y] (11 x 11) is summed per kBasisChunk paths on FP32 matrix cores or in FP64,
// the chunk sums in FP64 in chunk order; the normal equations (ridge 1e-12 n, plus kSaRidge x the
// diagonal for the swap-rate / annuity basis) are solved in FP64: elimination with partial
// pivoting for the state basis, Cholesky for the swap-rate / annuity basis.
constexpr int32_t kBasisState = CCR_BASIS_STATE, kBasisSwapRateAnnuity = CCR_BASIS_SWAP_RATE_ANNUITY;
constexpr int kBasisMax = 10;      // regression terms
constexpr int kBasisChunk = 4096;  // paths per workgroup of the moment and Gram passes (256 threads x 16)
constexpr int kBasisMom = 6;       // n, sum u1, sum u1^2, sum u2, sum u2^2, sum y
constexpr int kBasisGram = 66;     // upper triangle of the 11 x 11 Gram of [phi | y]
constexpr int kBasisStride = 16;   // per record: n, m1, sd1, m2, sd2, degree, coefficients[10]
constexpr double kSaRidge = 1e-4;  // swap-rate / annuity basis: ridge relative to the Gram diagonal
constexpr double kSaDegenerate = 1e-6;  // ... and a regressor with sd <= this x its rms is constant
struct RegRec {
  int32_t b;      // global Bermudan
  int32_t j;      // simulation date index on its set's grid
  int32_t k;      // exercise index (kind 0) or the next exercise (kind 1)
  int32_t kind;   // 0: exercise decision, 1: continuation
  int32_t ann, flt;  // swap-rate / annuity basis: TermRanges of the tail's annuity and float leg at t_j
  int32_t pad[2];
};

struct BermDev {
  int32_t set, ne, degree, cash;
  double position;
  int32_t ev_rec;     // first of ne exercise-value TermRanges
  int32_t date_rec;   // first of the set's T BermDate records (one per simulation date)
  int32_t eidx_off;   // first of ne exercise date indices
  int32_t ev_off;     // first of ne x P exercise values / strategy values (x P)
};

struct BermDate {
  int32_t next_k;     // first exercise strictly after t_j (ne: none left)
  int32_t has_swap;   // the underlying's value terms are present (physical settlement, exercisable)
  TermRange swap;     // value of the underlying swap at t_j
};

// One swaption that can be nonzero on a continuation date: its global index, its date record
// (continuation fit, date_rec + j), its exercise count and position, and a copy of that date's
// BermDate, so the valuation reads everything static about it from one record (the per-record
// chain of dependent scalar loads was the valuation kernel's bound).
struct LiveRec {
  int32_t b, rec, ne, next_k;
  int32_t has_swap, pad;
  double position;
  TermRange swap;
};
static_assert(sizeof(LiveRec) == 72, "LiveRec layout");

// One tile of the continuation GEMM: nd dates of one epoch (their continuation-date indices are
// tcd[d_off ..]) x nb swaptions (global indices gberm[col_off ..]) with the strategy column of
// each in the epoch (gcols[col_off ..]: exercise-value row, -1: no exercise left). The swaptions
// of a tile belong to netting sets valued on the same dates: one set, or in a book every set
// with the same valued dates (the design rows of a date are then shared by all of them).
struct GemmTile {
  int32_t d_off, nd, nb, col_off;
  int32_t pad[4];
};

struct SetDev {
  int64_t stream_stride;   // normals per path block: (T - 1) * nz
  int64_t rows_off;        // first element of this set's exposure rows
  int64_t partial_off;     // first element of this set's wave partials
  int32_t num_dates, num_rep;
  int32_t date_off;        // into the DateRec / date-real / step arrays
  int32_t bond_off, fixed_off, fixstore_off, poly_off;
  int32_t block_begin, num_blocks, wave_begin;
  int32_t ring_len, has_csa, sec_type, csa_asset;
  int32_t cpty, own, num_slots;
  int32_t berm_beg, berm_end;  // Bermudan swaptions of this set
  int32_t cd_base;             // first continuation date (one per valued date) of this set's rows
  int32_t opt_off;             // into the option-term table
  int64_t state_off;           // first element of this set's [date][path] states (AMC), -1 without
  int64_t dom_off;             // first element of its domestic normals [step][2][path block] (AMC), -1 without
  int64_t im_off;              // initial-margin proxy: first element of its |V_j - V_j-1| rows [date][path]
  int64_t im_part_off;         // and of its per-wave deflator sums [wave][date]
  int64_t tpart_off;           // per-trade statistics: first element of its wave partials
  int32_t num_groups;          // trades with their own value (per-trade plans), else 0
  int32_t tdate_off;           // first of its per-(date, trade) valuation records
  double thr_cpty, thr_own, mta, ia, keep;  // CSA terms (thr_own < 0 encodes a one-way CSA)
};

struct GlobalDev {
  int32_t num_paths, antithetic, nz, ncorr;
  int32_t nrates, nassets, ncredits, nsets;
  int32_t fx_of[kMaxRates];        // rate r >= 1: asset converting it (-1: none)
  int32_t asset_rate[kMaxAssets];  // asset a: foreign rate group (1..), 0 = deterministic carry
  int32_t credit_off[kMaxCredits]; // first bridge normal of credit c
  int32_t credit_sub[kMaxCredits];
  int32_t foreign_off;             // first foreign-integral normal
  int32_t max_ring, max_slots, pad;
  double spot[kMaxAssets];
  double cir_kappa[kMaxCredits], cir_theta[kMaxCredits], cir_xi[kMaxCredits], cir_y0[kMaxCredits];
  double chol[kMaxCorr * kMaxCorr];
};

}  // namespace ccrhip
