/*
 * ccr_hip.h -- C ABI of libccr_hip.so, the MI355X (gfx950) exposure engine.
 *
 * The library replaces the hot path of the `ccr` C++17 library (svr-net/zlatko):
 * correlated risk-factor simulation, per-scenario trade revaluation (including
 * American Monte Carlo), CSA netting and collateral, and the fused EE/PFE/EPE/CVA/DVA
 * reductions. Nothing C++ crosses this boundary: plain structs, pointers, sizes and
 * int status codes. The C++17 `ccr::` mirror in include/ccr_hip.hpp builds these
 * structs from model and trade objects and rethrows the status codes as the exception
 * types the reference uses.
 *
 * Reference interfaces replaced (paths relative to the reference repository root):
 *   ccr::ExposureEngine::run / reportingGrid / simulationGrid
 *       include/ccr/exposure/exposure_engine.hpp:48-66, src/exposure/exposure_engine.cpp:35-81
 *   ccr::gpu::compile -> FusedPlan, runFusedReference -> FusedOutput, summarise -> FusedResult
 *       include/ccr/gpu/fused_exposure.hpp:338-386 (the compile / backend / summarise split)
 *   the backend callback of app::analyseFused   src/app/analyses.cpp:2148-2177
 *   ccr::unilateralCva / bilateralCva / pathwiseCva   include/ccr/cva/cva.hpp:20-53
 *   ccr::BermudanSwaption::valueWithExercise   src/instruments/bermudan_swaption.cpp:49-146
 *   ccr::ai::fitMlp (training-loop conventions)   src/ai/surrogate.cpp:161-245
 *
 * Ownership: the caller owns every input and pre-sizes every output array; calls are
 * blocking. A context owns its device buffers, its HIP stream and its kernels' scratch;
 * one context per host thread. On failure no output is written and a message is left
 * for ccr_hip_last_error().
 */
#ifndef CCR_HIP_H
#define CCR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCR_HIP_ABI_VERSION 2

/* ------------------------------------------------------------------ status codes */
/* The C++ mirror maps them to the reference's exceptions:
 * INVALID_ARGUMENT -> std::invalid_argument, OUT_OF_RANGE -> std::out_of_range,
 * everything else -> std::runtime_error (exposure_engine.cpp:51,66; fused_exposure.cpp:219-247). */
enum ccr_status {
  CCR_OK = 0,
  CCR_ERR_INVALID_ARGUMENT = 1,
  CCR_ERR_OUT_OF_RANGE = 2,
  CCR_ERR_UNSUPPORTED = 3,   /* the setup exceeds what the kernels implement (cf. gpu::limitation) */
  CCR_ERR_OUT_OF_MEMORY = 4,
  CCR_ERR_DEVICE = 5,        /* HIP runtime or launch failure */
  CCR_ERR_RUNTIME = 6
};

/* ------------------------------------------------------------------ market and models */

/* ccr::YieldCurve: continuously compounded zero rates, linear in time, flat outside
 * (src/market/yield_curve.cpp:20-45). times strictly increasing, n >= 1. */
typedef struct ccr_yield_curve {
  const double* times;
  const double* zero_rates;
  int32_t size;
} ccr_yield_curve;

/* ccr::CreditCurve: piecewise-flat hazards on (pillar[i-1], pillar[i]] (src/market/credit_curve.cpp:23-41). */
typedef struct ccr_credit_curve {
  const double* pillars;
  const double* hazards;
  int32_t size;
} ccr_credit_curve;

/* ccr::HullWhite1F (G1++), include/ccr/models/hull_white.hpp:17-57. */
typedef struct ccr_hull_white {
  ccr_yield_curve curve;
  double mean_reversion;
  double volatility;
} ccr_hull_white;

/* ccr::LognormalAsset, include/ccr/models/lognormal_asset.hpp:17-41.
 * foreign_rate < 0: the carry is the deterministic `carry` curve (the reference model).
 * foreign_rate = f >= 0: the carry is the integrated short rate of foreign Hull-White f
 * (multi-currency extension; `carry` is then ignored). */
typedef struct ccr_lognormal_asset {
  double spot;
  double volatility;
  ccr_yield_curve carry;
  int32_t foreign_rate;
} ccr_lognormal_asset;

/* ccr::CirIntensity (CIR++), include/ccr/models/cir_intensity.hpp:22-53. */
typedef struct ccr_cir_intensity {
  ccr_credit_curve market;
  double kappa, theta, xi, y0;
  int32_t substeps;
} ccr_cir_intensity;

/* ccr::ScenarioGenerator, include/ccr/models/scenario_generator.hpp:64-84.
 * Correlated drivers are ordered [domestic rate, foreign rates..., assets..., credits...]
 * (the reference order [rate, assets..., credits...] when there is no foreign rate).
 * Normals drawn per step: [correlated drivers | domestic HW integral | CIR bridge normals
 * per credit | foreign HW integral per foreign rate]. correlation: row-major
 * num_factors^2, NULL for the identity. */
typedef struct ccr_scenario_generator {
  ccr_hull_white rates;
  int32_t num_foreign_rates;
  const ccr_hull_white* foreign_rates;
  int32_t num_assets;
  const ccr_lognormal_asset* assets;
  int32_t num_credits;
  const ccr_cir_intensity* credits;
  const double* correlation;
} ccr_scenario_generator;

/* ccr::SimulationConfig, include/ccr/models/scenario_generator.hpp:48-52. */
typedef struct ccr_simulation_config {
  int64_t num_paths;
  uint64_t seed;
  int32_t antithetic;
} ccr_simulation_config;

/* ------------------------------------------------------------------ trades */

enum ccr_trade_kind {
  CCR_TRADE_SWAP = 0,               /* ccr::InterestRateSwap (interest_rate_swap.hpp:12-66) */
  CCR_TRADE_ASSET_FORWARD = 1,      /* ccr::AssetForward (asset_forward.hpp:11-36) */
  CCR_TRADE_BERMUDAN_SWAPTION = 2,  /* ccr::BermudanSwaption (bermudan_swaption.hpp:26-62) */
  CCR_TRADE_EUROPEAN_OPTION = 3     /* ccr::EuropeanOption (european_option.hpp:18-41) */
};

enum ccr_option_type { CCR_OPTION_CALL = 0, CCR_OPTION_PUT = 1 };  /* ccr::OptionType (math.hpp:12) */

/* One trade. Swap (and the Bermudan's underlying): `schedule` T_0 < ... < T_n as in the
 * reference (both legs accrue on it). A swap with a different floating-leg schedule sets
 * float_schedule (its fixed leg accrues on `schedule`); NULL/0 means the reference's single
 * schedule. currency 0 = domestic; f + 1 = foreign Hull-White f, valued in domestic units at
 * the asset whose foreign_rate is f. Forward: asset, notional, strike (= fixed_rate),
 * maturity. Bermudan: exercise_times, cash_settled, position, regression_degree. European
 * option: asset (deterministic carry), option_type, notional (negative: short), strike
 * (= fixed_rate), expiry (= maturity). */
typedef struct ccr_trade {
  int32_t kind;
  int32_t pay_fixed;
  int32_t currency;
  int32_t asset;
  double notional;
  double fixed_rate;
  double maturity;
  const double* schedule;
  int32_t schedule_size;
  const double* float_schedule;
  int32_t float_schedule_size;
  const double* exercise_times;
  int32_t num_exercises;
  int32_t cash_settled;
  double position;
  int32_t regression_degree;
  int32_t option_type;
} ccr_trade;

enum ccr_collateral_type { CCR_COLLATERAL_CASH = 0, CCR_COLLATERAL_BOND = 1, CCR_COLLATERAL_ASSET = 2 };

/* ccr::CollateralAgreement, include/ccr/exposure/collateral.hpp:14-36. threshold_own = +inf
 * is a one-way CSA. */
typedef struct ccr_collateral {
  double threshold_counterparty;
  double threshold_own;
  double minimum_transfer_amount;
  double independent_amount;
  double margin_period_of_risk;
  int32_t type;
  double haircut;
  double bond_maturity;
  int32_t asset_index;
} ccr_collateral;

/* ccr::NettingSet (exposure_engine.hpp:19-23) plus the credit indices the fused pipeline
 * takes from FusedOptions::counterparty: counterparty_credit / own_credit index the
 * generator's credits (-1: none) for the pathwise CVA / DVA increments. */
typedef struct ccr_netting_set {
  int32_t num_trades;
  const ccr_trade* trades;
  int32_t has_collateral;
  ccr_collateral collateral;
  int32_t counterparty_credit;
  int32_t own_credit;
} ccr_netting_set;

/* ------------------------------------------------------------------ results */

/* ccr::ExposureProfile on the reporting dates (exposure_profile.hpp:18-42) plus the fused
 * pipeline's extra per-date statistics (fused_exposure.hpp:36-38, FusedResult). Every array
 * has `capacity` entries allocated by the caller (query the size with
 * ccr_hip_reporting_grid). NULL arrays are skipped. cva_increments[k] (k >= 1) is
 * E[1/2 (D V+_{k-1} + D V+_k)(Q_{k-1} - Q_k)] with Q the counterparty's pathwise survival;
 * dva_increments the same for -V- and the own credit. exposure_rows, if not NULL, receives
 * max(V - C, 0) as [reporting date][path] (capacity * num_paths doubles). */
typedef struct ccr_exposure_out {
  int32_t capacity;
  int32_t num_dates;  /* written */
  double* times;
  double* expected_value;
  double* expected_exposure;
  double* expected_negative_exposure;
  double* potential_future_exposure;
  double* discounted_expected_exposure;
  double* discounted_expected_negative_exposure;
  double* mean_deflator;
  double* mean_survival;
  double* cva_increments;
  double* dva_increments;
  double* exposure_rows;
} ccr_exposure_out;

/* ------------------------------------------------------------------ context */

enum ccr_precision {
  CCR_FP64 = 0,  /* FP64 everywhere: the parity mode (1e-10 relative to the FP64 oracle) */
  CCR_FP32 = 1,  /* FP32 path evolution (Kahan-compensated integrals of the short rate and the
                    log-spots) and FP32 revaluation: merged-bond power series (FP32 coefficients,
                    Horner in FP32), exponential terms, fixed coupons, options, collateral and the
                    Bermudans' exercise values; the American Monte Carlo regressions, the
                    statistics' wavefront sums and the final statistics stay FP64 */
  CCR_BF16 = 2   /* MLP surrogate only: BF16 matrix-core products, FP32 accumulation */
};

typedef struct ccr_hip_options {
  int32_t device;      /* HIP device ordinal */
  int32_t precision;   /* default ccr_precision of runs */
  int32_t block_size;  /* threads per workgroup of the path kernel (64..256, a multiple of 64); 0 = per
                          plan, the size with the most resident waves per CU (ccr_hip_autotune
                          measures the candidates instead) */
} ccr_hip_options;

typedef struct ccr_hip_ctx ccr_hip_ctx;
typedef struct ccr_hip_plan ccr_hip_plan;

/* Kernel timings of the last execute on the context's stream, from HIP events. */
typedef struct ccr_hip_timing {
  double path_ms;        /* path kernel (simulation, revaluation, CSA, per-block sums) */
  double reduce_ms;      /* fixed-order reduction of the per-block sums */
  double pfe_ms;         /* PFE order-statistic select */
  double total_ms;       /* first launch start to last launch end */
  double upload_ms;      /* host-to-device copies made by compile (draws, tables) */
  double draws_ms;       /* host generation of the normal stream (compile) */
  double amc_ms;         /* American Monte Carlo: states pass and regressions (before the path kernel) */
} ccr_hip_timing;

/* Shape of a compiled plan. */
typedef struct ccr_hip_plan_info {
  int32_t num_sets;
  int64_t num_paths;
  int32_t max_sim_dates;
  int32_t max_reporting_dates;
  int32_t num_normals;        /* normals per step */
  int64_t total_terms;        /* valuation terms over all sets and dates */
  int64_t stream_normals;     /* length of the replayed normal stream */
  int64_t revaluations;       /* sum over sets of trades x paths x reporting dates */
  int64_t term_evaluations;   /* sum over sets, paths and valued dates of the terms evaluated */
  int64_t hbm_bytes;          /* algorithmic HBM bytes of one execute: normals replayed by every
                                 set + exposure rows written and read back by the PFE select */
  int32_t precision;
  int32_t block_size;
  int32_t num_bermudans;      /* Bermudan swaptions valued by American Monte Carlo */
  int32_t num_exercises;      /* their exercise dates, summed */
  int32_t num_continuation_dates; /* (set, date) pairs with continuation regressions */
  int32_t max_exercises;
  int64_t normals_read;       /* normals the path kernel reads: every set replays its own slice of
                                 the stream, (paths / 2) x (T_s - 1) x nz per set with antithetics */
} ccr_hip_plan_info;

/* ------------------------------------------------------------------ MLP surrogate */

/* Multilayer perceptron: `inputs` -> hidden_layers x `hidden` (activation) -> `outputs` (linear),
 * as ccr::ai::Mlp (include/ccr/ai/surrogate.hpp:44-50; the reference has two tanh hidden
 * layers). Parameters are flattened layer by layer: W_l [fan_out][fan_in] row-major, then
 * b_l [fan_out]. Standardisers (ccr::ai::Standardiser, surrogate.hpp:29-34) are stored as
 * x_mean[inputs], x_sd[inputs], y_mean[outputs], y_sd[outputs]. */
enum ccr_activation { CCR_ACT_TANH = 0, CCR_ACT_SILU = 1 };
typedef struct ccr_mlp_desc {
  int32_t inputs;
  int32_t hidden;
  int32_t hidden_layers;
  int32_t outputs;
  int32_t activation;
} ccr_mlp_desc;

/* ccr::ai::MlpOptions (surrogate.hpp:51-56): full-batch Adam (beta 0.9 / 0.999, eps 1e-8,
 * bias-corrected, weight decay on weights only). precision: CCR_FP32 (vector units, any shape)
 * or CCR_BF16 (BF16 matrix cores with FP32 accumulation, hidden = 128). */
typedef struct ccr_mlp_train_options {
  int32_t epochs;
  double learning_rate;
  double weight_decay;
  uint32_t seed;
  int32_t precision;
  int32_t warm;            /* nonzero: start from the parameters and standardisers passed in */
  int32_t batch;           /* 0: full batch (the reference); else Adam steps over consecutive
                              mini-batches of this many rows (BF16 only, a multiple of 64) */
} ccr_mlp_train_options;

/* fitMlp (src/ai/surrogate.cpp:143-245) on the GPU: rows X [n][inputs], Y [n][outputs]. Unless
 * options->warm, fits the standardisers (into stdz) and initialises params from options->seed;
 * then options->epochs full-batch Adam steps. Writes the trained params, the loss (MSE on
 * standardised rows) after training, and the training time in milliseconds. */
int32_t ccr_hip_mlp_train(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, const ccr_mlp_train_options* options,
                          const double* X, const double* Y, int64_t n, double* params, double* stdz, double* loss,
                          double* train_ms);
/* Loss and its gradient with respect to params on standardised rows xs, ys (one forward and
 * one backward pass in the given precision: CCR_FP32 or CCR_BF16). */
int32_t ccr_hip_mlp_gradient(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, int32_t precision, const double* params,
                             const double* xs, const double* ys, int64_t n, double* grad, double* loss);
/* Mlp::predict (surrogate.cpp:157) for rows X [n][inputs] into Y [n][outputs]. */
int32_t ccr_hip_mlp_predict(ccr_hip_ctx* ctx, const ccr_mlp_desc* desc, int32_t precision, const double* params,
                            const double* stdz, const double* X, int64_t n, double* Y);

/* ------------------------------------------------------------------ exposure surrogate
 * For the trade type without a closed form (Bermudan swaptions): an MLP (BF16 matrix cores, SiLU)
 * maps the per-date state -- (t, x, running max and min of x along the path, x at every exercise
 * date already passed) -- to the Bermudan value of netting sets, one output per set. The netting
 * sets that share one simulated state (every set with Bermudans of a portfolio-book plan; a set
 * alone otherwise) share one network (at most 128 outputs each), so a book's networks are
 * evaluated once per (date, path) for all its sets. Trained on the pathwise values the American
 * Monte Carlo method produced on a training plan, then checked on an independent hold-out plan:
 * a set keeps the surrogate only if its hold-out EE stays within gate_threshold (relative to the
 * peak plain EE); otherwise that set falls back to the regression method (which then runs for
 * the plan). No counterpart in the reference. */
typedef struct ccr_hip_surrogate ccr_hip_surrogate;

typedef struct ccr_surrogate_options {
  int32_t train_paths;     /* paths of the training plan used as samples (0: all) */
  int32_t hidden;          /* 128 */
  int32_t hidden_layers;   /* e.g. 3 */
  int32_t epochs;
  double learning_rate;
  double weight_decay;
  uint32_t seed;
  double gate_threshold;   /* e.g. 0.005 */
  int32_t batch;           /* Adam mini-batch rows (a multiple of 64; 0: full batch); rows are shuffled */
} ccr_surrogate_options;

typedef struct ccr_surrogate_report {
  int32_t candidate_sets;  /* netting sets holding Bermudan swaptions */
  int32_t enabled_sets;    /* of which the gate kept the surrogate */
  int64_t samples;         /* training rows over all candidate sets */
  double train_ms;         /* GPU training time (all sets) */
  double loss;             /* largest final training loss (standardised MSE) */
  double max_ee_error;     /* worst hold-out EE error (relative to the peak plain EE) */
  int32_t networks;        /* networks trained (one per group of sets sharing a simulated state) */
} ccr_surrogate_report;

/* Trains on train_plan (executed plain) and gates on holdout_plan (executed plain and with the
 * surrogate). Both plans must hold the same netting sets (other paths / seeds). */
int32_t ccr_hip_surrogate_train(ccr_hip_ctx* ctx, ccr_hip_plan* train_plan, ccr_hip_plan* holdout_plan,
                                const ccr_surrogate_options* options, ccr_hip_surrogate** out,
                                ccr_surrogate_report* report);
/* Gate outcome for netting set `set`: enabled (1/0; -1: not a candidate) and hold-out EE error. */
int32_t ccr_hip_surrogate_set_info(const ccr_hip_surrogate* surrogate, int32_t set, int32_t* enabled,
                                   double* ee_error);
/* ccr_hip_execute / _async with the surrogate valuing the enabled sets' Bermudans. */
int32_t ccr_hip_execute_surrogate(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* surrogate,
                                  ccr_exposure_out* outs);
int32_t ccr_hip_execute_surrogate_async(ccr_hip_ctx* ctx, ccr_hip_plan* plan, const ccr_hip_surrogate* surrogate);
void ccr_hip_surrogate_destroy(ccr_hip_surrogate* surrogate);

int32_t ccr_hip_abi_version(void);

/* Creates a context on options->device (NULL: device 0, FP64, autotuned block). */
int32_t ccr_hip_create(const ccr_hip_options* options, ccr_hip_ctx** out);
void ccr_hip_destroy(ccr_hip_ctx* ctx);
/* Message of the last failed call on ctx ("" after a success); ctx NULL: the last
 * failure of ccr_hip_create on this thread. */
const char* ccr_hip_last_error(const ccr_hip_ctx* ctx);

/* Host-forced failure for tests of error handling: the next call that reaches the given
 * stage fails with the listed code, without touching the device. 0 clears. */
enum ccr_fault {
  CCR_FAULT_NONE = 0,
  CCR_FAULT_ALLOC = 1,    /* device allocation -> CCR_ERR_OUT_OF_MEMORY */
  CCR_FAULT_LAUNCH = 2,   /* kernel launch -> CCR_ERR_DEVICE */
  CCR_FAULT_PLAN = 3      /* plan validation -> CCR_ERR_INVALID_ARGUMENT */
};
int32_t ccr_hip_set_fault(ccr_hip_ctx* ctx, int32_t fault);

/* No counterpart in the reference (its engine materialises per netting set and frees on return).
   Plans compiled after ccr_hip_set_shared_workspace(ctx, 1) keep their per-execute buffers
   (exposure rows, wave partials, American Monte Carlo states, values and regressions) in one
   workspace of the context, sized to the largest such plan, instead of owning them: a portfolio
   split into waves of netting sets (one plan per wave) then needs HBM for one wave's buffers.
   The buffers of such a plan are valid from its execute until the next execute of any plan of
   the context (ccr_hip_execute copies the profiles out before returning); their plans cannot
   train a surrogate. 0 restores owned buffers for later compiles. */
int32_t ccr_hip_set_shared_workspace(ccr_hip_ctx* ctx, int32_t enable);

/* initialMarginProfile (src/xva/xva.cpp:59-76) of every netting set, on the GPU: plans compiled
 * after ccr_hip_set_initial_margin(ctx, quantile, mpor) with quantile in (0, 1] also write
 * |V(t_k) - V(t_k-1)| of the netted value on every simulation date and select its quantile over
 * the paths with the PFE kernel; quantile <= 0 switches it off for later compiles. After an
 * execute, ccr_hip_initial_margin gives one value per simulation date of netting set `set`
 * (ccr_hip_simulation_grid; 0 at t = 0): quantile x sqrt(mpor / dt) x the mean deflator. */
int32_t ccr_hip_set_initial_margin(ccr_hip_ctx* ctx, double quantile, double margin_period_of_risk);
int32_t ccr_hip_initial_margin(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set, double* margin,
                               int32_t capacity, int32_t* size);

/* Nested conditional initial margin (gpu::FusedOptions::imInnerPaths / imQuantile /
 * imMarginPeriod, include/ccr/gpu/fused_exposure.hpp:293-304; the kernel's inner Monte Carlo,
 * src/gpu/fused_kernels.cpp:445-479; the grid extension, src/gpu/fused_exposure.cpp:199-212).
 * Plans compiled after ccr_hip_set_nested_initial_margin(ctx, inner_paths, quantile, mpor) with
 * inner_paths >= 2 add t + mpor of every reporting date t (up to the horizon) to each netting
 * set's simulation grid and, on every execute, run from every path and reporting date t
 * inner_paths inner paths (antithetic pairs, from the path's state at t, drawn from a
 * counter-based generator keyed by the seed, netting set, path, date and pair) to t + mpor, and
 * take IM(t, path) = the quantile of |V(t + mpor) - V(t)| of the uncollateralised netted value
 * over them (0 where t + mpor is past the horizon). At most 1,024 inner paths, and the
 * quantile's order statistic must lie among the 16 largest samples (the reference's rule).
 * Netting sets with Bermudan swaptions, book plans and per-trade plans are not supported.
 * inner_paths = 0 switches it off for later compiles. After an execute,
 * ccr_hip_nested_initial_margin gives per reporting date of netting set `set`: E[D IM(t)]
 * (FusedResult::conditionalImDiscounted), E[IM(t)] (conditionalImMean) and the 5 / 50 / 95%
 * quantiles of IM(t) over the paths (conditionalImBands). */
typedef struct ccr_nested_im_out {
  int32_t capacity;        /* entries available in each array below */
  int32_t num_dates;       /* written: reporting dates */
  double* discounted;      /* E[D IM] */
  double* mean;            /* E[IM] */
  double* band05;          /* 5% quantile of IM over the paths */
  double* band50;
  double* band95;
} ccr_nested_im_out;
int32_t ccr_hip_set_nested_initial_margin(ccr_hip_ctx* ctx, int32_t inner_paths, double quantile,
                                          double margin_period_of_risk);
int32_t ccr_hip_nested_initial_margin(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set,
                                      ccr_nested_im_out* out);

/* Graph mode (no counterpart in the reference): while enabled, ccr_hip_execute and
 * ccr_hip_execute_async replay the plan's whole execute -- states pass, regressions, path kernel,
 * reductions, selects -- as one HIP graph, captured on the plan's first execute in this mode (and
 * again if its shared workspace has moved). Results are those of the eager launches; surrogate
 * executes stay eager. ccr_hip_last_timing then reports the whole replay in total_ms (the stage
 * fields are 0). */
int32_t ccr_hip_set_graph(ccr_hip_ctx* ctx, int32_t enable);

/* Portfolio book (src/app/analyses.cpp:1604-1709, 1727-1766): plans compiled after
 * ccr_hip_set_book(ctx, 1) simulate every netting set on the union of the sets' simulation grids
 * and report on the union of their reporting grids, from one seed, so rates, FX, intensities and
 * deflators are pathwise identical across the book and EE, ENE, EE*, ENE* and the pathwise CVA /
 * DVA increments of different netting sets add up exactly. */
int32_t ccr_hip_set_book(ccr_hip_ctx* ctx, int32_t enable);
/* The book's grids (src/app/analyses.cpp:1690-1709): the union of the sets' reporting grids and
 * the union of their simulation grids built on it. Writes up to the capacities and the full
 * sizes. Needs no device. */
int32_t ccr_hip_book_grids(const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                           double* reporting, int32_t reporting_capacity, int32_t* reporting_size, double* simulation,
                           int32_t simulation_capacity, int32_t* simulation_size);
/* A book compiled as several plans (waves of its netting sets): book plans compiled afterwards
 * report on `reporting` and simulate on `simulation` (the whole book's grids from
 * ccr_hip_book_grids) instead of the union of their own sets' grids, so every wave replays the
 * same scenarios. Compile fails with CCR_ERR_INVALID_ARGUMENT if a set needs a date missing from
 * them. NULL or size 0 clears. Sets with Bermudan swaptions share one states pass and one
 * regression design per date within a book plan. */
int32_t ccr_hip_set_book_grids(ccr_hip_ctx* ctx, const double* reporting, int32_t reporting_size,
                               const double* simulation, int32_t simulation_size);
/* Book plans compiled afterwards run the book kernel with this many netting sets per workgroup
 * (1, 2, 4 or 8; default 4; 1 = one set per workgroup): a lane steps its path's market once per
 * date and values every set of its group, sharing the fixing states. Results are bit-identical
 * for every choice. Plans with European options, per-trade statistics or the initial-margin
 * proxy run one set per workgroup. */
int32_t ccr_hip_set_book_group(ccr_hip_ctx* ctx, int32_t sets_per_workgroup);

/* Bermudan regressions (PolynomialRegression as BermudanSwaption::valueWithExercise uses it,
 * src/instruments/bermudan_swaption.cpp:77-93, 110-116; src/core/regression.cpp:29-90) for plans
 * compiled after the call. basis -1 (the default): the state basis 1, z, z^2, z^3 on per-date
 * designs shared by the swaptions of a date. Otherwise every exercise decision and every
 * continuation (swaption, valued date) is its own regression record:
 *   CCR_BASIS_STATE              the same state basis (reproduces the default's fits);
 *   CCR_BASIS_SWAP_RATE_ANNUITY  the 10 monomials of total degree <= 3 in the standardised swap
 *                                rate and annuity of the tail swap entered at the (next) exercise,
 *                                the swaption's degree capped at 3 as the total degree; no
 *                                counterpart in the reference, whose basis is the state alone.
 * gram_precision CCR_FP64 sums the Gram of [basis | target] in FP64; CCR_FP32 on FP32 matrix cores
 * (FP32 products and accumulation per 4,096 paths, FP64 across path chunks). The normal equations
 * (ridge 1e-12 n) are solved in FP64 by elimination with partial pivoting (matrix.cpp:70-96). */
enum ccr_regression_basis { CCR_BASIS_STATE = 0, CCR_BASIS_SWAP_RATE_ANNUITY = 1 };
int32_t ccr_hip_set_regression(ccr_hip_ctx* ctx, int32_t basis, int32_t gram_precision);
/* sumProfiles (src/app/analyses.cpp:1769-1784): the additive fields (expected value, EE, ENE, EE*,
 * ENE*, CVA and DVA increments) summed over the netting sets of counterparty credit
 * `counterparty` (-1: every set) of one book execute; times, mean deflator and survival are the
 * first set's; PFE is 0 (quantiles do not add). CCR_ERR_INVALID_ARGUMENT if the sets' grids differ,
 * CCR_ERR_OUT_OF_RANGE if no set has that counterparty. Needs no device. */
int32_t ccr_book_totals(const ccr_exposure_out* outs, const ccr_netting_set* sets, int32_t num_sets,
                        int32_t counterparty, ccr_exposure_out* total);

/* Per-trade statistics (gpu::FusedOptions::perTrade and FusedResult::TradeStats,
 * include/ccr/gpu/fused_exposure.hpp:251-256, 299; the estimators of marginalExposureContributions
 * and incrementalExposure, src/exposure/allocation.cpp:8-45): plans compiled after
 * ccr_hip_set_trade_statistics(ctx, 1) value every trade of a netting set on its own (the set's
 * value is their sum in trade order) and give, per trade and reporting date, the expectations
 * below, with V the set's value before collateral and D the deflator. At most 16 trades per
 * netting set and no Bermudan swaptions (CCR_ERR_UNSUPPORTED at compile). */
typedef struct ccr_trade_statistics {
  int32_t capacity;                /* doubles each non-null array can hold (>= reporting dates) */
  int32_t num_dates;               /* out: reporting dates written */
  double* marginal_discounted;     /* E[D V_i 1{V > 0}]: Euler contribution to EE* */
  double* marginal;                /* E[V_i 1{V > 0}]: Euler contribution to EE */
  double* incremental_discounted;  /* E[D (V+ - (V - V_i)+)]: EE* the trade adds */
  double* standalone_discounted;   /* E[D V_i+] */
  double* standalone;              /* E[V_i+] */
} ccr_trade_statistics;
int32_t ccr_hip_set_trade_statistics(ccr_hip_ctx* ctx, int32_t enable);
/* After an execute: the statistics of trade `trade` (its index in the set) of netting set `set`;
 * null arrays are skipped. */
int32_t ccr_hip_trade_statistics(ccr_hip_ctx* ctx, const ccr_hip_plan* plan, int32_t set, int32_t trade,
                                 ccr_trade_statistics* out);

/* ccr::NormalGenerator(seed).fill(out, n) (include/ccr/core/random.hpp:13-26): std::mt19937_64 +
 * std::normal_distribution<double>, the stream every plan replays. Generated on up to 16 host
 * threads (the engine run sequentially, the polar attempts in parallel chunks), bit-identical to
 * the sequential stream. Needs no device. */
int32_t ccr_hip_normals(uint64_t seed, int64_t n, double* out);

/* ExposureEngine::reportingGrid / simulationGrid of one netting set (exposure_engine.cpp:35-48).
 * Writes up to `capacity` times and the full size into *size. Needs no device. */
int32_t ccr_hip_reporting_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                               double* times, int32_t capacity, int32_t* size);
int32_t ccr_hip_simulation_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                double* times, int32_t capacity, int32_t* size);

/* gpu::compile analogue: builds the FP64 per-date tables of every netting set on the host,
 * generates the reference normal stream (std::mt19937_64 + std::normal_distribution<double>,
 * as ccr::NormalGenerator), and uploads tables and stream to the device. Each set is simulated
 * on its own grid exactly as ExposureEngine::run would (one generate() per set, same seed). */
int32_t ccr_hip_compile(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator,
                        const ccr_simulation_config* config, const double* base_grid, int32_t base_size,
                        const ccr_netting_set* sets, int32_t num_sets, double pfe_quantile, int32_t precision,
                        ccr_hip_plan** out);
int32_t ccr_hip_plan_info_get(const ccr_hip_plan* plan, ccr_hip_plan_info* info);
/* Runs the kernels of a compiled plan (inputs already resident) and writes one
 * ccr_exposure_out per netting set. */
int32_t ccr_hip_execute(ccr_hip_ctx* ctx, ccr_hip_plan* plan, ccr_exposure_out* outs);
/* Same, without reading results back (for timing loops). */
int32_t ccr_hip_execute_async(ccr_hip_ctx* ctx, ccr_hip_plan* plan);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation). ccr_hip_execute for a plan
 * compiled after ccr_hip_set_shared_workspace(ctx, 1), with its per-execute buffers (exposure rows,
 * wave partials, American Monte Carlo states, exercise values, regressions and valuation rows)
 * placed in a library-owned static device array of 64 MiB per GPU (held until the process ends)
 * instead of any allocation: the call makes no hipMalloc / hipFree and no host allocation except
 * the message of a failure. The statistics are copied straight into the caller's arrays.
 * Errors: CCR_ERR_INVALID_ARGUMENT for a plan compiled without the shared workspace,
 * CCR_ERR_OUT_OF_RANGE when its buffers exceed 64 MiB (e.g. ~50,000 paths x 120 dates for one
 * FP64 netting set), CCR_ERR_UNSUPPORTED for plans with the initial-margin proxy or per-trade
 * statistics and for FP32 exposure rows. Every call on one GPU shares that GPU's array, which keeps
 * the last call's contents (every kernel writes what it later reads, so nothing is cleared): calls
 * that overlap on one GPU -- two contexts, two host threads -- race. On a device error the output
 * arrays may be partly written (num_dates is set only on success). */
int32_t ccr_hip_execute_noalloc(ccr_hip_ctx* ctx, ccr_hip_plan* plan, ccr_exposure_out* outs);
/* Waits for the context's stream, then for the whole device. */
int32_t ccr_hip_synchronize(ccr_hip_ctx* ctx);
int32_t ccr_hip_last_timing(const ccr_hip_ctx* ctx, ccr_hip_timing* timing);
void ccr_hip_plan_destroy(ccr_hip_plan* plan);

/* ExposureEngine::run for several netting sets: compile + execute + free. */
int32_t ccr_hip_run_exposure(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator,
                             const ccr_simulation_config* config, const double* base_grid, int32_t base_size,
                             const ccr_netting_set* sets, int32_t num_sets, double pfe_quantile,
                             int32_t precision, ccr_exposure_out* outs);

/* Autotuner: compiles the workload with min(config->num_paths, sample_paths) paths for every
 * workgroup size of the path kernel (64, 128, 192, 256), times each execute (best of 3, HIP
 * events) and makes the fastest the context's workgroup size for later compiles. Writes the
 * chosen size and the four path-kernel times (ms). Results do not depend on the choice. */
int32_t ccr_hip_autotune(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                         const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                         int32_t precision, int64_t sample_paths, int32_t* block_size, double* times_ms);

/* Autotuner of the book kernel's shape (no counterpart in the reference): for a book context
 * (ccr_hip_set_book), compiles the netting sets with min(config->num_paths, sample_paths) paths
 * for every (workgroup size 64 / 128 / 192 / 256) x (netting sets per workgroup 1 / 2 / 4 / 8)
 * and times the path kernel of each (best of 3 executes, HIP events); the fastest pair becomes
 * the context's workgroup size and book group for later compiles. times_ms (16 entries, workgroup
 * size major; may be NULL) receives every candidate's time (a shape the kernel's LDS rules out
 * reads 0). Results do not depend on the choice. The sample should give every shape several
 * workgroups per compute unit (tens of thousands of paths): a small one favours small groups. */
int32_t ccr_hip_autotune_book(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator,
                              const ccr_simulation_config* config, const double* base_grid, int32_t base_size,
                              const ccr_netting_set* sets, int32_t num_sets, int32_t precision, int64_t sample_paths,
                              int32_t* block_size, int32_t* sets_per_workgroup, double* times_ms);

/* BermudanSwaption::valueWithExercise (src/instruments/bermudan_swaption.cpp:49-146) for one
 * Bermudan on its own scenario set: the simulation grid is `grid` merged with the trade's event
 * dates (as ExposureEngine::simulationGrid does), the regressions run on the GPU in `precision`.
 * Writes the price (mean of the time-0 value, times the position) and, if exercise_index is not
 * NULL, each path's first exercise index (-1: never exercised). */
int32_t ccr_hip_bermudan(ccr_hip_ctx* ctx, const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                         const double* grid, int32_t grid_size, const ccr_trade* trade, int32_t precision,
                         double* price, int32_t* exercise_index);

/* ------------------------------------------------------------------ profile statistics (FP64, host) */

/* ExposureProfile::expectedPositiveExposure (src/exposure/exposure_profile.cpp:52-65): the
 * right-point time average of EE over (0, horizon]. */
double ccr_expected_positive_exposure(const double* times, const double* ee, int32_t size, double horizon);
/* ExposureProfile::effectiveExpectedExposure (:67-75): the running maximum of EE, into eff[size]. */
void ccr_effective_expected_exposure(const double* ee, int32_t size, double* eff);
/* ExposureProfile::effectiveExpectedPositiveExposure (:77-79). */
double ccr_effective_expected_positive_exposure(const double* times, const double* ee, int32_t size, double horizon);

/* ------------------------------------------------------------------ CVA (FP64, host) */

/* unilateralCva (cva.cpp:22-33) and bilateralCva (cva.cpp:34-48) on a profile. */
double ccr_unilateral_cva(const double* times, const double* discounted_ee, int32_t size,
                          const ccr_credit_curve* counterparty, double recovery);
void ccr_bilateral_cva(const double* times, const double* discounted_ee, const double* discounted_ene,
                       int32_t size, const ccr_credit_curve* counterparty, double recovery_counterparty,
                       const ccr_credit_curve* own, double recovery_own, double* cva, double* dva);

/* ------------------------------------------------------------------ XVA on a profile (FP64, host) */

/* ccr::XvaParameters (include/ccr/xva/xva.hpp:12-21); ccr_xva_defaults fills the reference's
 * defaults (borrow 1%, lend 0, margin funding 1%, hurdle 10%, capital ratio 8%, risk weight 1,
 * alpha 1.4). */
typedef struct ccr_xva_parameters {
  double funding_spread_borrow, funding_spread_lend, margin_funding_spread;
  double hurdle_rate, capital_ratio, risk_weight, alpha;
} ccr_xva_parameters;
void ccr_xva_defaults(ccr_xva_parameters* parameters);

/* ccr::Xva (xva.hpp:33-42): every adjustment as a positive cost or benefit. */
typedef struct ccr_xva_result {
  double cva, dva, fca, fba, mva, kva;
} ccr_xva_result;

/* exposureXva (src/xva/xva.cpp:32-43): CVA, FCA and KVA of a discounted EE profile into
 * cva_fca_kva[3]. Returns CCR_OK, or CCR_ERR_INVALID_ARGUMENT for a null pointer. */
int32_t ccr_exposure_xva(const double* times, const double* discounted_ee, int32_t size,
                         const ccr_credit_curve* counterparty, double recovery_counterparty,
                         const ccr_credit_curve* own, const ccr_xva_parameters* parameters, double* cva_fca_kva);
/* allocateExposureXva (xva.cpp:97-106): CVA, FCA and KVA of each trade's discounted Euler
 * contribution (row i of contributions[num_trades][size], e.g. ccr_trade_statistics'
 * marginal_discounted) into out[num_trades][3]; the rows add up to the netting set's. */
int32_t ccr_allocate_exposure_xva(const double* times, const double* contributions, int32_t num_trades, int32_t size,
                                  const ccr_credit_curve* counterparty, double recovery_counterparty,
                                  const ccr_credit_curve* own, const ccr_xva_parameters* parameters, double* out);
/* fundingBenefit (xva.cpp:45-51) from the discounted ENE (<= 0). */
double ccr_funding_benefit(const double* times, const double* discounted_ene, int32_t size,
                           const ccr_credit_curve* counterparty, const ccr_credit_curve* own,
                           double funding_spread_lend);
/* marginValuationAdjustment (xva.cpp:53-57) from a discounted initial-margin profile. */
double ccr_margin_valuation_adjustment(const double* times, const double* discounted_initial_margin,
                                       int32_t size, const ccr_credit_curve* counterparty,
                                       const ccr_credit_curve* own, double margin_funding_spread);
/* xva (xva.cpp:78-96) of a profile; discounted_initial_margin may be NULL (MVA = 0). */
int32_t ccr_xva(const double* times, const double* discounted_ee, const double* discounted_ene, int32_t size,
                const double* discounted_initial_margin, const ccr_credit_curve* counterparty,
                double recovery_counterparty, const ccr_credit_curve* own, double recovery_own,
                const ccr_xva_parameters* parameters, ccr_xva_result* out);

#ifdef __cplusplus
}
#endif

#endif /* CCR_HIP_H */
