/*
 * ccr_oracle.h -- the FP64 CPU oracle of the exposure path. TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch restatement of the reference's CPU engine (ccr::ExposureEngine and what it
 * calls), used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker. The product (libccr_hip.so) never links, loads or calls it.
 *
 * It consumes the same C structs as the product's ABI (include/ccr_hip.h) and draws the same
 * normal stream (std::mt19937_64 + std::normal_distribution<double>, include/ccr/core/random.hpp:13-26
 * of the reference).
 */
#ifndef CCR_ORACLE_H
#define CCR_ORACLE_H

#include "../include/ccr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Wall-clock split of the last ccr_oracle_run_exposure on this thread (seconds). */
typedef struct ccr_oracle_timing {
  double draws_s;
  double simulate_s;  /* includes draws */
  double pricing_s;
  double collateral_s;
  double statistics_s;
  double total_s;
  int64_t revaluations; /* trades x paths x reporting dates */
} ccr_oracle_timing;

/* ExposureEngine::run for each netting set (src/exposure/exposure_engine.cpp:50-81). `threads`
 * parallelises pricing over paths (1 = the reference's single thread); results do not depend on
 * it. err receives a message on failure. Returns a ccr_status. */
int32_t ccr_oracle_run_exposure(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                const double* base_grid, int32_t base_size, const ccr_netting_set* sets,
                                int32_t num_sets, double pfe_quantile, int32_t threads, ccr_exposure_out* outs,
                                char* err, int32_t err_capacity);
int32_t ccr_oracle_last_timing(ccr_oracle_timing* out);
/* ccr_oracle_run_exposure with the normals replayed from `normals` (num_normals of them, in the
 * generator's draw order; every netting set's generate() starts at their beginning) instead of
 * the seeded stream. */
int32_t ccr_oracle_run_exposure_replay(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                       const double* base_grid, int32_t base_size, const ccr_netting_set* sets,
                                       int32_t num_sets, double pfe_quantile, int32_t threads, const double* normals,
                                       int64_t num_normals, ccr_exposure_out* outs, char* err, int32_t err_capacity);
/* ScenarioGenerator::generate on `grid` (seeded stream, or replayed normals when normals is not
 * NULL). Outputs [path][date] (foreign states, spots, survivals stacked per factor); NULL skips. */
int32_t ccr_oracle_scenarios(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* grid, int32_t grid_size, const double* normals, int64_t num_normals,
                             double* x, double* deflator, double* foreign_x, double* spots, double* survival,
                             char* err, int32_t err_capacity);

/* Grids (src/core/time_grid.cpp, src/exposure/exposure_engine.cpp:35-48). */
int32_t ccr_oracle_time_grid(const double* times, int32_t n, double* out, int32_t capacity, int32_t* size);
int32_t ccr_oracle_uniform_grid(double horizon, int32_t steps, double* out, int32_t capacity, int32_t* size);
int32_t ccr_oracle_lagged_grid(const double* grid, int32_t n, double lag, double* out, int32_t capacity,
                               int32_t* size);
int32_t ccr_oracle_index_at_or_before(const double* grid, int32_t n, double t);
int32_t ccr_oracle_reporting_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                  double* out, int32_t capacity, int32_t* size);
int32_t ccr_oracle_simulation_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                   double* out, int32_t capacity, int32_t* size);

/* The normal stream of ccr::NormalGenerator(seed): n draws. */
void ccr_oracle_normals(uint64_t seed, int64_t n, double* out);
/* Raw std::mt19937_64(seed) outputs (for the standard's known answer). */
void ccr_oracle_mt19937_64(uint64_t seed, int64_t n, uint64_t* out);

/* Model pieces, for known-answer tests. */
double ccr_oracle_discount(const ccr_yield_curve* curve, double t);
double ccr_oracle_instantaneous_forward(const ccr_yield_curve* curve, double t);
double ccr_oracle_survival(const ccr_credit_curve* curve, double t);
double ccr_oracle_zero_bond(const ccr_hull_white* hw, double t, double T, double x);
void ccr_oracle_hw_evolve(const ccr_hull_white* hw, double t0, double t1, double x0, double z1, double z2,
                          double drift, double* x1, double* integrated_rate);
double ccr_oracle_cir_survival(const ccr_cir_intensity* cir, double t);
double ccr_oracle_quantile(const double* values, int64_t n, double p);
int32_t ccr_oracle_cholesky(const double* a, int32_t n, double* l);
int32_t ccr_oracle_solve(const double* a, const double* b, int32_t n, double* x);
/* PolynomialRegression::fit on masked samples (mask NULL = all); coefficients (degree+1
 * entries, the unused ones 0), center, scale and the effective degree. */
/* Basis of the Bermudan regressions of later runs: -1 / 0 the reference's state basis
 * (bermudan_swaption.cpp:77-93), 1 the swap-rate / annuity basis (total degree <= 3 in the
 * standardised swap rate and annuity of the tail swap; no reference counterpart). */
/* Nested conditional initial margin of netting set set_index (per reporting date: E[D IM], E[IM],
 * 5 / 50 / 95% quantiles); sample_path >= 0: the inner samples of (sample_path, sample_rep). */
int32_t ccr_oracle_nested_im(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t set_index,
                             int32_t inner_paths, double quantile, double mpor, int32_t threads, double* discounted,
                             double* mean, double* band05, double* band50, double* band95, int32_t capacity,
                             int32_t* num_dates, int64_t sample_path, int32_t sample_rep, double* samples,
                             double* sample_im, char* err, int32_t err_capacity);
int32_t ccr_oracle_set_regression_basis(int32_t basis);
/* Swap rate (P(t,T_0) - P(t,T_n)) / A and annuity A = sum_i tau_i P(t,T_i) of a schedule at (t, x). */
int32_t ccr_oracle_swap_rate_annuity(const ccr_hull_white* hw, const double* sched, int32_t n, double t, double x,
                                     double* rate, double* annuity);
/* The swap-rate / annuity regression on given regressors: coefficients[10] (zeros past the
 * effective terms) and params = mean / sd of each regressor, effective total degree. */
int32_t ccr_oracle_sa_regression(const double* u1, const double* u2, const double* y, const uint8_t* mask, int64_t n,
                                 int32_t degree, double* coefficients, double* params);
int32_t ccr_oracle_regression(const double* x, const double* y, const uint8_t* mask, int64_t n, int32_t degree,
                              double* coefficients, double* center, double* scale, int32_t* effective_degree);
/* simulateCollateral on a [paths][dates] value matrix (unit_price NULL for cash). */
int32_t ccr_oracle_collateral(const ccr_collateral* csa, const double* values, const double* grid, int32_t paths,
                              int32_t dates, const double* unit_price, double* out);
/* ExposureProfile::compute on [paths][dates] values and deflators. */
int32_t ccr_oracle_profile(const double* values, const double* deflator, int32_t paths, int32_t dates,
                           double pfe_quantile, double* expected_value, double* ee, double* ene, double* pfe,
                           double* dee, double* dene);
/* ExposureProfile::expectedPositiveExposure / effectiveExpectedPositiveExposure. */
double ccr_oracle_epe(const double* times, const double* ee, int32_t n, double horizon);
double ccr_oracle_effective_epe(const double* times, const double* ee, int32_t n, double horizon);
/* unilateralCva / bilateralCva (src/cva/cva.cpp:11-48). */
double ccr_oracle_unilateral_cva(const double* times, const double* dee, int32_t n, const ccr_credit_curve* cpty,
                                 double recovery);
void ccr_oracle_bilateral_cva(const double* times, const double* dee, const double* dene, int32_t n,
                              const ccr_credit_curve* cpty, double rc, const ccr_credit_curve* own, double ro,
                              double* cva, double* dva);
/* Bermudan valuation of one trade on its own on a generated scenario set: price and per-path
 * exercise index (-1 = never). */
int32_t ccr_oracle_bermudan(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                            const double* grid, int32_t grid_size, const ccr_trade* trade, double* price,
                            int32_t* exercise_index);

#ifdef __cplusplus
}
#endif

#endif
