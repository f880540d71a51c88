// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY. A thin extern "C" driver over the reference
// library itself, compiled (by oracle/Makefile) against the reference's headers and sources where
// they lie, into oracle/_ref/libccr_ref.so. Tests use it to pin the FP64 oracle against the real
// reference engine; nothing in the product loads it.
//
// It builds reference objects from the C structs of include/ccr_hip.h and calls
// ccr::ExposureEngine::run (src/exposure/exposure_engine.cpp:50-81), ccr::pathwiseCva
// (src/cva/cva.cpp:55-77), ccr::initialMarginProfile and ccr::xva (src/xva/xva.cpp:59-96) and
// ccr::BermudanSwaption::valueWithExercise, and the fused pipeline's CPU mirror
// (gpu::compile / runFusedReference / summarise) for its nested conditional initial margin. Only reference-
// representable inputs are accepted: no foreign rates; a two-leg swap is expressed as
// IRS(fixed schedule, K) + IRS(float schedule, 0) - IRS(fixed schedule, 0), all reference swaps.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/ccr_hip.h"
#include "ccr/ai/surrogate.hpp"
#include "ccr/core/math.hpp"
#include "ccr/core/random.hpp"
#include "ccr/cva/cva.hpp"
#include "ccr/exposure/allocation.hpp"
#include "ccr/exposure/exposure_engine.hpp"
#include "ccr/gpu/fused_exposure.hpp"
#include "ccr/instruments/asset_forward.hpp"
#include "ccr/instruments/bermudan_swaption.hpp"
#include "ccr/instruments/credit_default_swap.hpp"
#include "ccr/instruments/european_option.hpp"
#include "ccr/instruments/interest_rate_swap.hpp"
#include "ccr/xva/xva.hpp"

namespace {

std::shared_ptr<const ccr::YieldCurve> curveOf(const ccr_yield_curve& c) {
  return std::make_shared<ccr::YieldCurve>(std::vector<double>(c.times, c.times + c.size),
                                           std::vector<double>(c.zero_rates, c.zero_rates + c.size));
}

std::shared_ptr<const ccr::CreditCurve> creditOf(const ccr_credit_curve& c) {
  return std::make_shared<ccr::CreditCurve>(std::vector<double>(c.pillars, c.pillars + c.size),
                                            std::vector<double>(c.hazards, c.hazards + c.size));
}

ccr::ScenarioGenerator generatorOf(const ccr_scenario_generator& g) {
  if (g.num_foreign_rates != 0) throw std::invalid_argument("ref driver: the reference has no foreign rates");
  auto hw = std::make_shared<ccr::HullWhite1F>(curveOf(g.rates.curve), g.rates.mean_reversion, g.rates.volatility);
  std::vector<std::shared_ptr<const ccr::LognormalAsset>> assets;
  for (int a = 0; a < g.num_assets; ++a) {
    if (g.assets[a].foreign_rate >= 0) throw std::invalid_argument("ref driver: deterministic carry only");
    assets.push_back(std::make_shared<ccr::LognormalAsset>("A" + std::to_string(a), g.assets[a].spot,
                                                           g.assets[a].volatility, curveOf(g.assets[a].carry)));
  }
  std::vector<std::shared_ptr<const ccr::CirIntensity>> credits;
  for (int c = 0; c < g.num_credits; ++c) {
    const auto& m = g.credits[c];
    credits.push_back(std::make_shared<ccr::CirIntensity>("C" + std::to_string(c), creditOf(m.market), m.kappa,
                                                          m.theta, m.xi, m.y0, static_cast<std::size_t>(m.substeps)));
  }
  const std::size_t n = 1 + assets.size() + credits.size();
  ccr::Matrix corr;
  if (g.correlation) {
    corr = ccr::Matrix(n, n);
    for (std::size_t i = 0; i < n; ++i)
      for (std::size_t j = 0; j < n; ++j) corr(i, j) = g.correlation[i * n + j];
  }
  return ccr::ScenarioGenerator(hw, assets, credits, corr);
}

ccr::InterestRateSwap swapOf(const std::string& id, const double* s, int n, double notional, double k, bool pay) {
  return ccr::InterestRateSwap(id, notional, k, std::vector<double>(s, s + n),
                               pay ? ccr::InterestRateSwap::Direction::PayFixed
                                   : ccr::InterestRateSwap::Direction::ReceiveFixed);
}

void addTrades(const ccr_trade& d, int i, ccr::NettingSet& ns) {
  const std::string id = "T" + std::to_string(i);
  if (d.kind == CCR_TRADE_SWAP) {
    if (d.currency != 0) throw std::invalid_argument("ref driver: domestic swaps only");
    const bool pay = d.pay_fixed != 0;
    if (d.float_schedule && d.float_schedule_size > 0) {
      ns.trades.push_back(std::make_shared<ccr::InterestRateSwap>(
          swapOf(id + "a", d.schedule, d.schedule_size, d.notional, d.fixed_rate, pay)));
      ns.trades.push_back(std::make_shared<ccr::InterestRateSwap>(
          swapOf(id + "b", d.float_schedule, d.float_schedule_size, d.notional, 0.0, pay)));
      ns.trades.push_back(std::make_shared<ccr::InterestRateSwap>(
          swapOf(id + "c", d.schedule, d.schedule_size, d.notional, 0.0, !pay)));
    } else {
      ns.trades.push_back(std::make_shared<ccr::InterestRateSwap>(
          swapOf(id, d.schedule, d.schedule_size, d.notional, d.fixed_rate, pay)));
    }
  } else if (d.kind == CCR_TRADE_ASSET_FORWARD) {
    ns.trades.push_back(
        std::make_shared<ccr::AssetForward>(id, "A" + std::to_string(d.asset), d.notional, d.fixed_rate, d.maturity));
  } else if (d.kind == CCR_TRADE_EUROPEAN_OPTION) {
    ns.trades.push_back(std::make_shared<ccr::EuropeanOption>(
        id, "A" + std::to_string(d.asset), d.option_type == CCR_OPTION_CALL ? ccr::OptionType::Call : ccr::OptionType::Put,
        d.notional, d.fixed_rate, d.maturity));
  } else if (d.kind == CCR_TRADE_BERMUDAN_SWAPTION) {
    ns.trades.push_back(std::make_shared<ccr::BermudanSwaption>(
        id, swapOf(id + "u", d.schedule, d.schedule_size, d.notional, d.fixed_rate, d.pay_fixed != 0),
        std::vector<double>(d.exercise_times, d.exercise_times + d.num_exercises),
        d.cash_settled ? ccr::Settlement::Cash : ccr::Settlement::Physical, d.position,
        static_cast<std::size_t>(d.regression_degree)));
  } else {
    throw std::invalid_argument("ref driver: unknown trade kind");
  }
}

int fail(char* err, int cap, const char* what, int code) {
  if (err && cap > 0) std::snprintf(err, static_cast<std::size_t>(cap), "%s", what);
  return code;
}

ccr::NettingSet nettingSetOf(const ccr_netting_set& set) {
  ccr::NettingSet ns;
  ns.id = "NS";
  for (int i = 0; i < set.num_trades; ++i) addTrades(set.trades[i], i, ns);
  if (set.has_collateral) {
    const ccr_collateral& c = set.collateral;
    ccr::CollateralAgreement csa;
    csa.thresholdCounterparty = c.threshold_counterparty;
    csa.thresholdOwn = c.threshold_own;
    csa.minimumTransferAmount = c.minimum_transfer_amount;
    csa.independentAmount = c.independent_amount;
    csa.marginPeriodOfRisk = c.margin_period_of_risk;
    csa.type = c.type == CCR_COLLATERAL_BOND    ? ccr::CollateralAgreement::Type::Bond
               : c.type == CCR_COLLATERAL_ASSET ? ccr::CollateralAgreement::Type::Asset
                                                : ccr::CollateralAgreement::Type::Cash;
    csa.haircut = c.haircut;
    csa.bondMaturity = c.bond_maturity;
    csa.assetIndex = static_cast<std::size_t>(c.asset_index);
    ns.collateral = csa;
  }
  return ns;
}

}  // namespace

extern "C" {

// ccr::initialMarginProfile of the netted value (before collateral) of one netting set on its
// simulation grid, after ExposureEngine::run.
int32_t ccr_ref_initial_margin(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                               const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                               double quantile_level, double mpor, double* margin, int32_t capacity, int32_t* size,
                               char* err, int32_t err_capacity) {
  try {
    const ccr::ScenarioGenerator gen = generatorOf(*generator);
    const ccr::SimulationConfig cfg{static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0};
    const ccr::ExposureEngine engine(gen, cfg, ccr::TimeGrid(std::vector<double>(base_grid, base_grid + base_size)));
    const ccr::ExposureResult r = engine.run(nettingSetOf(*set), 0.95);
    const std::vector<double> im =
        ccr::initialMarginProfile(r.nettedValue, r.scenarios.grid, r.scenarios.deflator, quantile_level, mpor);
    if (size) *size = static_cast<int32_t>(im.size());
    if (capacity < static_cast<int32_t>(im.size())) return fail(err, err_capacity, "capacity", CCR_ERR_INVALID_ARGUMENT);
    for (std::size_t k = 0; k < im.size(); ++k) margin[k] = im[k];
    return CCR_OK;
  } catch (const std::exception& e) {
    return fail(err, err_capacity, e.what(), CCR_ERR_RUNTIME);
  }
}

// Per-trade statistics of one netting set on its reporting dates, stats[(i * 5 + m) * capacity + k]
// for trade i: m = 0 / 1 marginalExposureContributions with / without the deflator, 2
// incrementalExposure of trade i added to the other trades' net (with the deflator), 3 / 4 the
// trade's own discounted / undiscounted EE (ExposureProfile::compute of its values).
int32_t ccr_ref_trade_statistics(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                 const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                 double* stats, int32_t capacity, char* err, int32_t err_capacity) {
  try {
    const ccr::ScenarioGenerator gen = generatorOf(*generator);
    const ccr::SimulationConfig cfg{static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0};
    const ccr::ExposureEngine engine(gen, cfg, ccr::TimeGrid(std::vector<double>(base_grid, base_grid + base_size)));
    const ccr::ExposureResult r = engine.run(nettingSetOf(*set), 0.95);
    const std::vector<std::size_t>& idx = r.reportingIndices;
    const std::size_t nP = r.scenarios.numPaths, nR = idx.size();
    if (capacity < static_cast<int32_t>(nR)) return fail(err, err_capacity, "capacity", CCR_ERR_INVALID_ARGUMENT);
    auto columns = [&](const ccr::Matrix& m) {
      ccr::Matrix out(nP, nR, 0.0);
      for (std::size_t p = 0; p < nP; ++p)
        for (std::size_t k = 0; k < nR; ++k) out(p, k) = m(p, idx[k]);
      return out;
    };
    const ccr::Matrix net = columns(r.nettedValue), deflator = columns(r.scenarios.deflator);
    std::vector<ccr::Matrix> values;
    for (const ccr::Matrix& m : r.tradeValues) values.push_back(columns(m));
    const auto discounted = ccr::marginalExposureContributions(values, net, &deflator);
    const auto plain = ccr::marginalExposureContributions(values, net, nullptr);
    std::vector<double> times;
    for (std::size_t k = 0; k < nR; ++k) times.push_back(r.scenarios.grid[idx[k]]);
    const ccr::TimeGrid grid(times);
    for (std::size_t i = 0; i < values.size(); ++i) {
      ccr::Matrix rest = net;
      for (std::size_t p = 0; p < nP; ++p)
        for (std::size_t k = 0; k < nR; ++k) rest(p, k) = net(p, k) - values[i](p, k);
      const auto inc = ccr::incrementalExposure(rest, values[i], &deflator);
      const ccr::ExposureProfile own = ccr::ExposureProfile::compute(values[i], grid, deflator);
      const std::vector<double>* rows[5] = {&discounted[i], &plain[i], &inc, &own.discountedExpectedExposure,
                                            &own.expectedExposure};
      for (int m = 0; m < 5; ++m)
        for (std::size_t k = 0; k < nR; ++k) stats[(i * 5 + static_cast<std::size_t>(m)) * capacity + k] = (*rows[m])[k];
    }
    return CCR_OK;
  } catch (const std::exception& e) {
    return fail(err, err_capacity, e.what(), CCR_ERR_RUNTIME);
  }
}

// ccr::xva of a profile given by its times, discounted EE and discounted ENE.
int32_t ccr_ref_xva(const double* times, const double* discounted_ee, const double* discounted_ene, int32_t size,
                    const double* discounted_im, const ccr_credit_curve* counterparty, double recovery_counterparty,
                    const ccr_credit_curve* own, double recovery_own, const ccr_xva_parameters* parameters,
                    ccr_xva_result* out) {
  try {
    ccr::ExposureProfile profile;
    profile.times.assign(times, times + size);
    profile.discountedExpectedExposure.assign(discounted_ee, discounted_ee + size);
    profile.discountedExpectedNegativeExposure.assign(discounted_ene, discounted_ene + size);
    ccr::XvaParameters p;
    p.fundingSpreadBorrow = parameters->funding_spread_borrow;
    p.fundingSpreadLend = parameters->funding_spread_lend;
    p.marginFundingSpread = parameters->margin_funding_spread;
    p.hurdleRate = parameters->hurdle_rate;
    p.capitalRatio = parameters->capital_ratio;
    p.riskWeight = parameters->risk_weight;
    p.alpha = parameters->alpha;
    const std::vector<double> im = discounted_im ? std::vector<double>(discounted_im, discounted_im + size)
                                                 : std::vector<double>();
    const ccr::Xva x = ccr::xva(profile, im, *creditOf(*counterparty), recovery_counterparty, *creditOf(*own),
                                recovery_own, p);
    *out = ccr_xva_result{x.cva, x.dva, x.fca, x.fba, x.mva, x.kva};
    return CCR_OK;
  } catch (const std::exception&) {
    return CCR_ERR_RUNTIME;
  }
}

int32_t ccr_ref_run_exposure(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                             double pfe_quantile, ccr_exposure_out* out, double* pathwise_cva, double* pathwise_dva,
                             char* err, int32_t err_capacity) {
  try {
    const ccr::ScenarioGenerator gen = generatorOf(*generator);
    const ccr::SimulationConfig cfg{static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0};
    const ccr::ExposureEngine engine(gen, cfg, ccr::TimeGrid(std::vector<double>(base_grid, base_grid + base_size)));
    const ccr::ExposureResult r = engine.run(nettingSetOf(*set), pfe_quantile);
    const auto& pr = r.profile;
    const std::size_t nR = pr.times.size();
    if (out->capacity < static_cast<int32_t>(nR)) return fail(err, err_capacity, "capacity", CCR_ERR_INVALID_ARGUMENT);
    out->num_dates = static_cast<int32_t>(nR);
    for (std::size_t k = 0; k < nR; ++k) {
      const std::size_t j = r.reportingIndices[k];
      if (out->times) out->times[k] = pr.times[k];
      if (out->expected_value) out->expected_value[k] = pr.expectedValue[k];
      if (out->expected_exposure) out->expected_exposure[k] = pr.expectedExposure[k];
      if (out->expected_negative_exposure) out->expected_negative_exposure[k] = pr.expectedNegativeExposure[k];
      if (out->potential_future_exposure) out->potential_future_exposure[k] = pr.potentialFutureExposure[k];
      if (out->discounted_expected_exposure) out->discounted_expected_exposure[k] = pr.discountedExpectedExposure[k];
      if (out->discounted_expected_negative_exposure)
        out->discounted_expected_negative_exposure[k] = pr.discountedExpectedNegativeExposure[k];
      if (out->mean_deflator) {
        double m = 0.0;
        for (std::size_t p = 0; p < r.scenarios.numPaths; ++p) m += r.scenarios.deflator(p, j);
        out->mean_deflator[k] = m / static_cast<double>(r.scenarios.numPaths);
      }
      if (out->exposure_rows)
        for (std::size_t p = 0; p < r.scenarios.numPaths; ++p)
          out->exposure_rows[k * r.scenarios.numPaths + p] = std::max(r.exposureValue(p, j), 0.0);
    }
    if (pathwise_cva)
      *pathwise_cva = set->counterparty_credit >= 0
                          ? ccr::pathwiseCva(r.scenarios, r.exposureValue, static_cast<std::size_t>(set->counterparty_credit),
                                             0.0, r.reportingIndices)
                          : 0.0;
    if (pathwise_dva) {
      // The reference's pathwise CVA applied to -V and the own credit (the same estimator).
      *pathwise_dva = 0.0;
      if (set->own_credit >= 0) {
        ccr::Matrix neg = r.exposureValue * -1.0;
        *pathwise_dva =
            ccr::pathwiseCva(r.scenarios, neg, static_cast<std::size_t>(set->own_credit), 0.0, r.reportingIndices);
      }
    }
    return CCR_OK;
  } catch (const std::exception& e) {
    return fail(err, err_capacity, e.what(), CCR_ERR_RUNTIME);
  }
}

// The nested conditional initial margin of the reference's fused pipeline on its CPU mirror
// (gpu::FusedOptions::imInnerPaths, src/gpu/fused_reference.cpp:371-402): per reporting date
// FusedResult::conditionalImDiscounted and conditionalImMean (f32 kernels, its own counter-based
// draws).
int32_t ccr_ref_fused_nested_im(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                int32_t inner_paths, double quantile, double mpor, double* discounted, double* mean,
                                int32_t capacity, int32_t* num_dates, char* err, int32_t err_capacity) {
  try {
    const ccr::ScenarioGenerator gen = generatorOf(*generator);
    const ccr::SimulationConfig cfg{static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0};
    const ccr::ExposureEngine engine(gen, cfg, ccr::TimeGrid(std::vector<double>(base_grid, base_grid + base_size)));
    ccr::gpu::FusedOptions o;
    o.imInnerPaths = static_cast<std::size_t>(inner_paths);
    o.imQuantile = quantile;
    o.imMarginPeriod = mpor;
    const ccr::gpu::FusedPlan plan = ccr::gpu::compile(engine, nettingSetOf(*set), o);
    const ccr::gpu::FusedResult r = ccr::gpu::summarise(plan, ccr::gpu::runFusedReference(plan));
    const std::size_t n = r.conditionalImMean.size();
    if (num_dates) *num_dates = static_cast<int32_t>(n);
    if (static_cast<int32_t>(n) > capacity) return fail(err, err_capacity, "capacity", CCR_ERR_INVALID_ARGUMENT);
    for (std::size_t k = 0; k < n; ++k) {
      if (discounted) discounted[k] = r.conditionalImDiscounted[k];
      if (mean) mean[k] = r.conditionalImMean[k];
    }
    return CCR_OK;
  } catch (const std::exception& e) {
    return fail(err, err_capacity, e.what(), CCR_ERR_RUNTIME);
  }
}

int32_t ccr_ref_bermudan(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                         const double* grid, int32_t grid_size, const ccr_trade* trade, double* price,
                         int32_t* exercise_index) {
  try {
    const ccr::ScenarioGenerator gen = generatorOf(*generator);
    const ccr::SimulationConfig cfg{static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0};
    const auto s = gen.generate(ccr::TimeGrid(std::vector<double>(grid, grid + grid_size)), cfg);
    ccr::NettingSet ns;
    addTrades(*trade, 0, ns);
    const auto* b = dynamic_cast<const ccr::BermudanSwaption*>(ns.trades[0].get());
    const auto v = b->valueWithExercise(s);
    *price = v.price;
    if (exercise_index)
      for (std::size_t p = 0; p < s.numPaths; ++p)
        exercise_index[p] = v.exerciseIndex[p] == ccr::BermudanSwaption::npos ? -1 : static_cast<int32_t>(v.exerciseIndex[p]);
    return CCR_OK;
  } catch (const std::exception&) {
    return CCR_ERR_RUNTIME;
  }
}

void ccr_ref_normals(uint64_t seed, int64_t n, double* out) {
  ccr::NormalGenerator g(seed);
  g.fill(out, static_cast<std::size_t>(n));
}

double ccr_ref_quantile(const double* v, int64_t n, double p) { return ccr::quantile(std::vector<double>(v, v + n), p); }

// The market parsing of the front door, on the reference's own functions: bootstrapCreditCurve
// (quarterly premiums) of CDS quotes on a yield curve into hazards[n] (pillars = sorted
// maturities), InterestRateSwap::vanilla(...).parRate, TimeGrid::standardExposureGrid and
// cvaRunningSpread's risky annuity.
int32_t ccr_ref_bootstrap_credit(const ccr_yield_curve* curve, const double* maturities, const double* spreads,
                                 int32_t n, double recovery, double* pillars, double* hazards) {
  try {
    std::vector<ccr::CdsQuote> q;
    for (int32_t i = 0; i < n; ++i) q.push_back({maturities[i], spreads[i]});
    const ccr::CreditCurve cc = ccr::bootstrapCreditCurve(*curveOf(*curve), q, recovery);
    for (std::size_t i = 0; i < cc.pillars().size(); ++i) {
      pillars[i] = cc.pillars()[i];
      hazards[i] = cc.hazards()[i];
    }
    return CCR_OK;
  } catch (const std::exception&) {
    return CCR_ERR_INVALID_ARGUMENT;
  }
}
double ccr_ref_par_rate(const ccr_yield_curve* curve, double start, double tenor, int32_t freq) {
  const auto swap = ccr::InterestRateSwap::vanilla("s", 1.0, 0.0, start, tenor, freq,
                                                   ccr::InterestRateSwap::Direction::PayFixed);
  return swap.parRate(*curveOf(*curve));
}
int32_t ccr_ref_standard_grid(double horizon, double* out, int32_t capacity) {
  const auto g = ccr::TimeGrid::standardExposureGrid(horizon).times();
  for (std::size_t i = 0; i < g.size() && static_cast<int32_t>(i) < capacity; ++i) out[i] = g[i];
  return static_cast<int32_t>(g.size());
}
double ccr_ref_risky_annuity(const ccr_yield_curve* curve, const ccr_credit_curve* credit, double maturity) {
  return ccr::riskyAnnuity(*curveOf(*curve), *creditOf(*credit), maturity);
}

// ccr::ai::fitMlp on rows X [n][din], Y [n][dout]; writes W0, B0, W1, B1, W2, B2 (row-major) and
// the standardisers x_mean, x_sd, y_mean, y_sd.
int32_t ccr_ref_fit_mlp(const double* X, const double* Y, int64_t n, int32_t din, int32_t dout, int32_t hidden,
                        int32_t epochs, double lr, double wd, uint32_t seed, double* params, double* stdz) {
  try {
    ccr::ai::Rows rx, ry;
    for (int64_t r = 0; r < n; ++r) {
      rx.emplace_back(X + r * din, X + (r + 1) * din);
      ry.emplace_back(Y + r * dout, Y + (r + 1) * dout);
    }
    ccr::ai::MlpOptions o;
    o.hidden = static_cast<std::size_t>(hidden);
    o.epochs = static_cast<std::size_t>(epochs);
    o.learningRate = lr;
    o.weightDecay = wd;
    o.seed = seed;
    const ccr::ai::Mlp m = ccr::ai::fitMlp(rx, ry, o);
    for (std::size_t l = 0; l < 3; ++l) {
      for (const auto& row : m.W[l])
        for (double w : row) *params++ = w;
      for (double b : m.B[l]) *params++ = b;
    }
    for (double v : m.sx.mean) *stdz++ = v;
    for (double v : m.sx.sd) *stdz++ = v;
    for (double v : m.sy.mean) *stdz++ = v;
    for (double v : m.sy.sd) *stdz++ = v;
    return CCR_OK;
  } catch (const std::exception&) {
    return CCR_ERR_RUNTIME;
  }
}

}  // extern "C"
