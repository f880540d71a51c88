// plan.hpp -- host plan builder: turns a scenario generator, a simulation config and netting sets
// into the flat per-date tables of layout.hpp (the analogue of ccr::gpu::compile,
// src/gpu/fused_exposure.cpp:214-707, re-derived for one path per lane and FP64 parity).
#pragma once

#include <cstdint>
#include <vector>

#include "layout.hpp"
#include "model.hpp"

namespace ccrhip {

struct HostPlan {
  GlobalDev g{};
  std::vector<SetDev> sets;
  std::vector<double> steps;      // [sum T_s][kStepStride]
  std::vector<DateRec> dates;     // [sum T_s]
  std::vector<double> date_reals; // [sum T_s][kDateRealStride]
  std::vector<double> bonds;      // [n][2]
  std::vector<double> fixed;      // [n][4]
  std::vector<double> poly;       // bond-polynomial coefficients, ascending powers per (date, rate)
  std::vector<double> opts;       // European option terms [n][kOptStride]
  std::vector<int32_t> fixstore;  // [n][2] (rate, slot)
  std::vector<int32_t> block_set; // set of each workgroup (before the XCD remap)
  std::vector<std::vector<double>> rep_times;  // reporting dates per set
  std::vector<std::vector<double>> sim_times;  // simulation dates per set
  int64_t stream_len = 0;         // normals in the replayed stream
  int32_t max_dates = 0, max_rep = 0, total_waves = 0;
  int32_t block = 256;
  int32_t price_rows = 0;         // LDS rows of the collateral price ring (securities CSA only)
  int32_t precision = CCR_FP64;
  double pfe_q = 0.95;
  int64_t revaluations = 0;       // trades x paths x reporting dates
  int64_t term_evaluations = 0;   // exponential terms x paths over valued dates
  int64_t poly_evaluations = 0;   // polynomial coefficients x paths over valued dates
  int64_t total_terms = 0;
  // American Monte Carlo.
  std::vector<BermDev> berms;
  std::vector<BermDate> bdates;     // per Bermudan, one per simulation date of its set
  std::vector<TermRange> evterms;   // per Bermudan exercise: the tail swap at the exercise date
  std::vector<int32_t> eidx;        // exercise date indices
  int64_t state_len = 0;            // elements of the x / D state arrays
  int64_t dom_len = 0;              // domestic normals of the AMC states pass (both per step and path block)
  int64_t ev_len = 0;               // elements of the exercise-value / strategy arrays
  int32_t max_exercises = 0;
  int64_t cont_len = 0;             // kContStride x sum over Bermudans of their set's dates
  std::vector<int32_t> exlist;      // per exercise (global order): its Bermudan
  std::vector<int32_t> cdates;      // continuation dates: (set, date) pairs
  std::vector<double> ctimes;       // time of each continuation date
  std::vector<int32_t> bdate_berm;  // per BermDate record: its Bermudan
  std::vector<int32_t> bdate_cd;    // per BermDate record: its continuation date (-1: none)
  std::vector<LiveRec> clive;       // per continuation date: the swaptions with a value
  std::vector<int32_t> clive_rng;   // [begin, end) into clive per continuation date
  std::vector<GemmTile> tiles;      // continuation GEMM tiles
  std::vector<GemmTile> ntiles;     // narrow continuation GEMM tiles (groups of few Bermudans)
  std::vector<int32_t> gcols;       // strategy column per (tile, swaption), -1: none
  std::vector<int32_t> gberm;       // global Bermudan index per (tile, swaption)
  std::vector<int32_t> tcd;         // continuation date per (tile, date)
  // Continuation dates whose state column is another's (book plans: every set's date j has the
  // same x and D): cmap[cd] is the continuation date holding the moments and Gram of cd; cuniq
  // lists those (cmap[cd] == cd).
  std::vector<int32_t> cmap, cuniq;
  int32_t states_set = -1;          // book plans: the one set whose blocks run the states pass
  // Regression engine with a basis choice (basis >= 0; -1: the shared-design state-basis path).
  int32_t basis = -1;
  std::vector<RegRec> rrecs;        // exercise records first (per Bermudan, exercise 0..ne-1), then continuation
  std::vector<int32_t> rex;         // [Bermudan][max_exercises]: record of exercise k (-1: none)
  std::vector<int32_t> rcont;       // per BermDate record: its continuation record (-1: none)
  std::vector<TermRange> saterms;   // annuity / float-leg terms of the swap-rate / annuity records
  int32_t nex_recs = 0;             // exercise records (the continuation records follow)
  int32_t group = 0;                // book plans: netting sets per workgroup of the book kernel (0: one)
  // Per-trade plans: one valuation record (terms of that trade alone) per (set, date, trade),
  // [tdate_off + j G + i], and its reals; the set's own records then carry no terms.
  std::vector<DateRec> tdates;
  std::vector<double> tdreals;
  int32_t max_groups = 0;
  // Nested conditional initial margin (nested_inner >= 2): per simulation date of every set
  // ([date_off + j]), the date index of t_j + mpor for a reporting date j (-1: past the horizon,
  // or not a reporting date).
  int32_t nested_inner = 0;
  std::vector<int32_t> nested_target;
};


// ExposureEngine::reportingGrid / simulationGrid (src/exposure/exposure_engine.cpp:35-48).
TimeGrid reportingGrid(const TimeGrid& base, const ccr_netting_set& set);
TimeGrid simulationGrid(const TimeGrid& base, const ccr_netting_set& set);

// book: every set on the book's union grids (bookReportingGrid / bookSimulationGrid of the plan's
// sets, or book_rep / book_sim when given -- the grids of a whole book compiled as several plans,
// which must hold every set's own dates).
HostPlan buildPlan(const ccr_scenario_generator& gen, const ccr_simulation_config& cfg, const double* base_grid,
                   int32_t base_size, const ccr_netting_set* sets, int32_t num_sets, double pfe_quantile,
                   int32_t precision, int32_t block, bool per_trade = false, bool book = false,
                   bool value_all = false, const std::vector<double>* book_rep = nullptr,
                   const std::vector<double>* book_sim = nullptr, int32_t basis = -1, int32_t nested_inner = 0,
                   double nested_mpor = 0.0);

// The portfolio book's grids (src/app/analyses.cpp:1690-1709): the union of the sets' reporting
// grids, and the union of their simulation grids built on that reporting union.
TimeGrid bookReportingGrid(const TimeGrid& base, const ccr_netting_set* sets, int32_t num_sets);
TimeGrid bookSimulationGrid(const TimeGrid& rep, const ccr_netting_set* sets, int32_t num_sets);

// Workgroup size of the path kernel: the workgroups of every set, their waves and the offsets of
// the per-wave partial sums.
void assignBlocks(HostPlan& plan, int block);

// Power-series coefficients of one date's merged bond terms, truncated at `tolerance` (see
// layout.hpp, kPolyTolerance).
bool bondPolynomial(const double* terms, int n, double c0, double xmax, std::vector<double>& coef,
                    double tolerance = kPolyTolerance);

// The reference normal stream: std::mt19937_64(seed) + std::normal_distribution<double>, as
// ccr::NormalGenerator (include/ccr/core/random.hpp:13-26), n draws.
void generateStream(uint64_t seed, int64_t n, double* out);

}  // namespace ccrhip
