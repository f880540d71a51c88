// ccr_hip_reference.hpp -- the binding a ccr maintainer adds on the reference side. It is compiled
// inside a ccr build (it includes the reference's headers) and maps the reference's objects onto
// the C structs of ccr_hip.h:
//   ccr::ScenarioGenerator / ccr::SimulationConfig (include/ccr/models/scenario_generator.hpp:48-84)
//     -> ccr_scenario_generator / ccr_simulation_config
//   ccr::NettingSet with its trades and CSA (include/ccr/exposure/exposure_engine.hpp:19-23,
//     include/ccr/instruments/*.hpp, include/ccr/exposure/collateral.hpp:14-36)
//     -> ccr_netting_set / ccr_trade / ccr_collateral
// and runs ccr::ExposureEngine::run's job (exposure_engine.hpp:48-66) through a C entry, returning
// a ccr::ExposureResult (exposure_engine.hpp:26-37) whose profile, reporting indices and scenario
// grid / models are filled. The streamed GPU pipeline materialises no cube: rateState, deflator,
// assetValues, intensityState, survival, tradeValues, nettedValue, collateral and exposureValue stay
// empty (ccr_exposure_out.exposure_rows holds max(exposureValue, 0) on the reporting dates).
//
//   ccr::hip::Context ctx(0);                                  // ccr_hip.hpp
//   ccr::ExposureResult r = ccr::hip::run(ctx.get(), engine, nettingSet, 0.95);
//
// Trade types without a counterpart in ccr_hip.h (e.g. ccr::CreditDefaultSwap) raise
// std::invalid_argument, as do assets referenced by a name the generator does not hold.
#pragma once

#include <cmath>
#include <cstdint>
#include <deque>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ccr/exposure/exposure_engine.hpp"
#include "ccr/instruments/asset_forward.hpp"
#include "ccr/instruments/bermudan_swaption.hpp"
#include "ccr/instruments/european_option.hpp"
#include "ccr/instruments/interest_rate_swap.hpp"
#include "ccr_hip.h"

namespace ccr {
namespace hip {

// The C structs of one generator / configuration / base grid and of the netting sets added to
// it. Owns every array the structs point into: they stay valid while this object lives.
class ReferenceStructs {
 public:
  ReferenceStructs(const ::ccr::ScenarioGenerator& generator, const ::ccr::SimulationConfig& config,
                   const ::ccr::TimeGrid& baseGrid)
      : base_(baseGrid.times()) {
    gen_ = ccr_scenario_generator{};
    gen_.rates = hullWhite(*generator.rateModel());
    for (const auto& a : generator.assets()) {
      ccr_lognormal_asset c{};
      c.spot = a->spot();
      c.volatility = a->volatility();
      c.carry = curve(a->carryCurve());
      c.foreign_rate = -1;  // the reference's deterministic carry
      assets_.push_back(c);
      assetNames_.push_back(a->name());
    }
    for (const auto& m : generator.credits()) {
      ccr_cir_intensity c{};
      c.market = credit(m->marketCurve());
      c.kappa = m->kappa();
      c.theta = m->theta();
      c.xi = m->xi();
      c.y0 = m->y0();
      c.substeps = static_cast<int32_t>(m->substeps());
      credits_.push_back(c);
      creditNames_.push_back(m->name());
    }
    gen_.num_assets = static_cast<int32_t>(assets_.size());
    gen_.assets = assets_.empty() ? nullptr : assets_.data();
    gen_.num_credits = static_cast<int32_t>(credits_.size());
    gen_.credits = credits_.empty() ? nullptr : credits_.data();
    const ::ccr::Matrix& rho = generator.correlation();
    if (rho.rows() > 0) {
      std::vector<double>& c = keep(std::vector<double>(rho.rows() * rho.cols()));
      for (std::size_t i = 0; i < rho.rows(); ++i)
        for (std::size_t j = 0; j < rho.cols(); ++j) c[i * rho.cols() + j] = rho(i, j);
      gen_.correlation = c.data();
    }
    cfg_ = ccr_simulation_config{static_cast<int64_t>(config.numPaths), config.seed, config.antithetic ? 1 : 0};
  }
  ReferenceStructs(const ReferenceStructs&) = delete;
  ReferenceStructs& operator=(const ReferenceStructs&) = delete;

  // The struct of `ns` (valid while this object lives). counterparty / own credit: indices into
  // the generator's credits for the pathwise CVA / DVA increments (-1: none; creditIndex maps a
  // credit's name to its index).
  const ccr_netting_set& add(const ::ccr::NettingSet& ns, int32_t counterpartyCredit = -1, int32_t ownCredit = -1) {
    std::vector<ccr_trade>& trades = tradeLists_.emplace_back();
    for (const auto& t : ns.trades) trades.push_back(trade(*t));
    ccr_netting_set s{};
    s.num_trades = static_cast<int32_t>(trades.size());
    s.trades = trades.empty() ? nullptr : trades.data();
    s.counterparty_credit = counterpartyCredit;
    s.own_credit = ownCredit;
    if (ns.collateral) {
      const ::ccr::CollateralAgreement& a = *ns.collateral;
      s.has_collateral = 1;
      s.collateral.threshold_counterparty = a.thresholdCounterparty;
      s.collateral.threshold_own = a.thresholdOwn;
      s.collateral.minimum_transfer_amount = a.minimumTransferAmount;
      s.collateral.independent_amount = a.independentAmount;
      s.collateral.margin_period_of_risk = a.marginPeriodOfRisk;
      using Type = ::ccr::CollateralAgreement::Type;
      s.collateral.type = a.type == Type::Cash ? CCR_COLLATERAL_CASH
                          : a.type == Type::Bond ? CCR_COLLATERAL_BOND
                                                 : CCR_COLLATERAL_ASSET;
      s.collateral.haircut = a.haircut;
      s.collateral.bond_maturity = a.bondMaturity;
      s.collateral.asset_index = static_cast<int32_t>(a.assetIndex);
    }
    return sets_.emplace_back(s);
  }
  int32_t creditIndex(const std::string& name) const {
    for (std::size_t i = 0; i < creditNames_.size(); ++i)
      if (creditNames_[i] == name) return static_cast<int32_t>(i);
    throw std::invalid_argument("ccr_hip binding: no credit named " + name);
  }

  const ccr_scenario_generator& generator() const { return gen_; }
  const ccr_simulation_config& config() const { return cfg_; }
  const std::vector<double>& baseGrid() const { return base_; }

 private:
  std::vector<double>& keep(std::vector<double> v) { return arrays_.emplace_back(std::move(v)); }
  ccr_yield_curve curve(const ::ccr::YieldCurve& c) {
    const std::vector<double>& t = keep(c.times());
    const std::vector<double>& r = keep(c.zeroRates());
    return ccr_yield_curve{t.data(), r.data(), static_cast<int32_t>(t.size())};
  }
  ccr_credit_curve credit(const ::ccr::CreditCurve& c) {
    const std::vector<double>& p = keep(c.pillars());
    const std::vector<double>& h = keep(c.hazards());
    return ccr_credit_curve{p.data(), h.data(), static_cast<int32_t>(p.size())};
  }
  ccr_hull_white hullWhite(const ::ccr::HullWhite1F& m) {
    return ccr_hull_white{curve(m.curve()), m.meanReversion(), m.volatility()};
  }
  int32_t assetIndex(const std::string& name) const {
    for (std::size_t i = 0; i < assetNames_.size(); ++i)
      if (assetNames_[i] == name) return static_cast<int32_t>(i);
    throw std::invalid_argument("ccr_hip binding: no asset named " + name);
  }
  void swapLeg(const ::ccr::InterestRateSwap& s, ccr_trade& c) {
    c.pay_fixed = s.direction() == ::ccr::InterestRateSwap::Direction::PayFixed ? 1 : 0;
    c.currency = 0;
    c.notional = s.notional();
    c.fixed_rate = s.fixedRate();
    c.maturity = s.maturity();
    const std::vector<double>& sched = keep(s.schedule());
    c.schedule = sched.data();
    c.schedule_size = static_cast<int32_t>(sched.size());
  }
  ccr_trade trade(const ::ccr::Trade& t) {
    ccr_trade c{};
    c.position = 1.0;
    if (const auto* s = dynamic_cast<const ::ccr::InterestRateSwap*>(&t)) {
      c.kind = CCR_TRADE_SWAP;
      swapLeg(*s, c);
    } else if (const auto* f = dynamic_cast<const ::ccr::AssetForward*>(&t)) {
      c.kind = CCR_TRADE_ASSET_FORWARD;
      c.asset = assetIndex(f->asset());
      c.notional = f->notional();
      c.fixed_rate = f->strike();
      c.maturity = f->maturity();
    } else if (const auto* o = dynamic_cast<const ::ccr::EuropeanOption*>(&t)) {
      c.kind = CCR_TRADE_EUROPEAN_OPTION;
      c.asset = assetIndex(o->asset());
      c.option_type = o->type() == ::ccr::OptionType::Call ? CCR_OPTION_CALL : CCR_OPTION_PUT;
      c.notional = o->notional();
      c.fixed_rate = o->strike();
      c.maturity = o->expiry();
    } else if (const auto* b = dynamic_cast<const ::ccr::BermudanSwaption*>(&t)) {
      c.kind = CCR_TRADE_BERMUDAN_SWAPTION;
      swapLeg(b->underlying(), c);
      const std::vector<double>& ex = keep(b->exerciseTimes());
      c.exercise_times = ex.data();
      c.num_exercises = static_cast<int32_t>(ex.size());
      c.cash_settled = b->settlement() == ::ccr::Settlement::Cash ? 1 : 0;
      c.position = b->position();
      c.regression_degree = static_cast<int32_t>(b->regressionDegree());
    } else {
      throw std::invalid_argument("ccr_hip binding: trade " + t.id() + " has no counterpart in ccr_hip.h");
    }
    return c;
  }

  std::deque<std::vector<double>> arrays_;
  std::vector<ccr_lognormal_asset> assets_;
  std::vector<ccr_cir_intensity> credits_;
  std::vector<std::string> assetNames_, creditNames_;
  std::deque<std::vector<ccr_trade>> tradeLists_;
  std::deque<ccr_netting_set> sets_;
  ccr_scenario_generator gen_{};
  ccr_simulation_config cfg_{};
  std::vector<double> base_;
};

// The C entries the binding calls: an exposure run of one netting set, and the reporting /
// simulation grids (ExposureEngine::reportingGrid / simulationGrid). ccr_hip's by default;
// any implementation of the same C structs plugs in (the tests route it to a CPU checker).
struct EntryPoints {
  std::function<int32_t(const ccr_scenario_generator*, const ccr_simulation_config*, const double*, int32_t,
                        const ccr_netting_set*, double, ccr_exposure_out*)>
      run;
  std::function<int32_t(const double*, int32_t, const ccr_netting_set*, double*, int32_t, int32_t*)> reportingGrid;
  std::function<int32_t(const double*, int32_t, const ccr_netting_set*, double*, int32_t, int32_t*)> simulationGrid;
  std::function<std::string()> lastError;
};

inline EntryPoints hipEntryPoints(ccr_hip_ctx* ctx, int32_t precision = CCR_FP64) {
  EntryPoints e;
  e.run = [ctx, precision](const ccr_scenario_generator* g, const ccr_simulation_config* c, const double* base,
                           int32_t nbase, const ccr_netting_set* s, double q, ccr_exposure_out* out) {
    return ccr_hip_run_exposure(ctx, g, c, base, nbase, s, 1, q, precision, out);
  };
  e.reportingGrid = ccr_hip_reporting_grid;
  e.simulationGrid = ccr_hip_simulation_grid;
  e.lastError = [ctx]() { return std::string(ccr_hip_last_error(ctx) ? ccr_hip_last_error(ctx) : "ccr_hip error"); };
  return e;
}

namespace detail {
inline void check(int32_t code, const EntryPoints& e) {
  if (code == CCR_OK) return;
  const std::string m = e.lastError ? e.lastError() : std::string("ccr_hip error");
  if (code == CCR_ERR_INVALID_ARGUMENT) throw std::invalid_argument(m);
  if (code == CCR_ERR_OUT_OF_RANGE) throw std::out_of_range(m);
  throw std::runtime_error(m);
}
inline std::vector<double> grid(const EntryPoints& e,
                                const std::function<int32_t(const double*, int32_t, const ccr_netting_set*, double*,
                                                            int32_t, int32_t*)>& f,
                                const ReferenceStructs& s, const ccr_netting_set& set) {
  int32_t n = 0;
  check(f(s.baseGrid().data(), static_cast<int32_t>(s.baseGrid().size()), &set, nullptr, 0, &n), e);
  std::vector<double> t(static_cast<std::size_t>(n));
  check(f(s.baseGrid().data(), static_cast<int32_t>(s.baseGrid().size()), &set, t.data(), n, &n), e);
  return t;
}
}  // namespace detail

// ExposureEngine::run(nettingSet, pfeQuantile) through the entries `e`.
inline ::ccr::ExposureResult run(const EntryPoints& e, const ::ccr::ExposureEngine& engine, const ::ccr::NettingSet& nettingSet,
                          double pfeQuantile = 0.95) {
  ReferenceStructs s(engine.generator(), engine.config(), engine.baseGrid());
  const ccr_netting_set& set = s.add(nettingSet);
  const std::vector<double> rep = detail::grid(e, e.reportingGrid, s, set);
  const std::vector<double> sim = detail::grid(e, e.simulationGrid, s, set);
  ::ccr::ExposureResult r;
  ::ccr::ExposureProfile& p = r.profile;
  const std::size_t n = rep.size();
  for (auto* v : {&p.times, &p.expectedValue, &p.expectedExposure, &p.expectedNegativeExposure,
                  &p.potentialFutureExposure, &p.discountedExpectedExposure, &p.discountedExpectedNegativeExposure})
    v->assign(n, 0.0);
  ccr_exposure_out out{};
  out.capacity = static_cast<int32_t>(n);
  out.times = p.times.data();
  out.expected_value = p.expectedValue.data();
  out.expected_exposure = p.expectedExposure.data();
  out.expected_negative_exposure = p.expectedNegativeExposure.data();
  out.potential_future_exposure = p.potentialFutureExposure.data();
  out.discounted_expected_exposure = p.discountedExpectedExposure.data();
  out.discounted_expected_negative_exposure = p.discountedExpectedNegativeExposure.data();
  detail::check(e.run(&s.generator(), &s.config(), s.baseGrid().data(), static_cast<int32_t>(s.baseGrid().size()),
                      &set, pfeQuantile, &out),
                e);
  p.pfeQuantile = pfeQuantile;
  // The scenario set's grid and models; its state cubes are not materialised.
  r.scenarios.grid = ::ccr::TimeGrid(sim);
  r.scenarios.numPaths = engine.config().numPaths;
  r.scenarios.rateModel = engine.generator().rateModel();
  r.scenarios.assetModels = engine.generator().assets();
  r.scenarios.creditModels = engine.generator().credits();
  // Reporting dates as indices into the simulation grid (the grids hold the same doubles).
  for (std::size_t i = 0, j = 0; i < n; ++i) {
    while (j < sim.size() && sim[j] != rep[i]) ++j;
    if (j == sim.size()) throw std::runtime_error("ccr_hip binding: reporting date not on the simulation grid");
    r.reportingIndices.push_back(j);
  }
  return r;
}

// The same on the GPU: ccr_hip_run_exposure on `ctx` in the given precision.
inline ::ccr::ExposureResult run(ccr_hip_ctx* ctx, const ::ccr::ExposureEngine& engine, const ::ccr::NettingSet& nettingSet,
                          double pfeQuantile = 0.95, int32_t precision = CCR_FP64) {
  return run(hipEntryPoints(ctx, precision), engine, nettingSet, pfeQuantile);
}

}  // namespace hip
}  // namespace ccr
