// TEST INFRASTRUCTURE ONLY (needs the reference tree; built by `make -C oracle binding`).
//
// The reference-side binding (include/ccr_hip_reference.hpp) maps ccr:: objects onto the C structs
// of ccr_hip.h. This program builds reference portfolios as ccr:: objects, runs the reference's own
// ccr::ExposureEngine::run (src/exposure/exposure_engine.cpp:50-81) on them, and runs the same
// engine call through the binding with its entries routed to the FP64 oracle's C entry (which
// takes the same structs). Identical profiles, reporting indices and scenario grids show that the
// binding hands the C entry exactly the reference's inputs. It also checks that a trade type
// without a C counterpart is rejected with std::invalid_argument.
//
// Prints one line per case and "binding: all cases bit-identical" on success; exits 1 otherwise.
#include <cmath>
#include <cstdio>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ccr_hip_reference.hpp"
#include "../../oracle/ccr_oracle.h"
#include "ccr/instruments/asset_forward.hpp"
#include "ccr/instruments/bermudan_swaption.hpp"
#include "ccr/instruments/european_option.hpp"
#include "ccr/instruments/interest_rate_swap.hpp"

namespace {

using ccr::InterestRateSwap;

ccr::hip::EntryPoints oracleEntries() {
  ccr::hip::EntryPoints e;
  auto err = std::make_shared<std::string>();
  e.run = [err](const ccr_scenario_generator* g, const ccr_simulation_config* c, const double* base, int32_t nbase,
                const ccr_netting_set* s, double q, ccr_exposure_out* out) {
    char msg[512] = {0};
    const int32_t rc = ccr_oracle_run_exposure(g, c, base, nbase, s, 1, q, 2, out, msg, sizeof msg);
    *err = msg;
    return rc;
  };
  e.reportingGrid = ccr_oracle_reporting_grid;
  e.simulationGrid = ccr_oracle_simulation_grid;
  e.lastError = [err]() { return *err; };
  return e;
}

std::shared_ptr<const ccr::YieldCurve> curve(std::vector<double> t, std::vector<double> r) {
  return std::make_shared<ccr::YieldCurve>(std::move(t), std::move(r));
}

ccr::ScenarioGenerator generator(bool credits) {
  auto hw = std::make_shared<ccr::HullWhite1F>(curve({0.5, 2.0, 5.0, 10.0}, {0.020, 0.025, 0.030, 0.033}), 0.05, 0.01);
  std::vector<std::shared_ptr<const ccr::LognormalAsset>> assets{
      std::make_shared<ccr::LognormalAsset>("EURUSD", 1.1, 0.12, curve({1.0}, {0.01})),
      std::make_shared<ccr::LognormalAsset>("SPX", 100.0, 0.2, curve({1.0, 5.0}, {0.015, 0.02}))};
  std::vector<std::shared_ptr<const ccr::CirIntensity>> cirs;
  ccr::Matrix rho;
  if (credits) {
    cirs.push_back(std::make_shared<ccr::CirIntensity>(
        "CPTY", std::make_shared<ccr::CreditCurve>(std::vector<double>{1.0, 3.0, 5.0}, std::vector<double>{0.01, 0.015, 0.02}),
        0.5, 0.012, 0.08, 0.008, 4));
    rho = ccr::Matrix(4, 4, 0.0);
    for (int i = 0; i < 4; ++i) rho(i, i) = 1.0;
    rho(0, 1) = rho(1, 0) = 0.3;
    rho(1, 3) = rho(3, 1) = 0.25;
    rho(0, 2) = rho(2, 0) = -0.1;
  }
  return ccr::ScenarioGenerator(hw, assets, cirs, rho);
}

std::vector<double> schedule(double start, double end, double step) {
  std::vector<double> s;
  for (double t = start; t <= end + 1e-12; t += step) s.push_back(t);
  return s;
}

ccr::NettingSet portfolio(int variant) {
  ccr::NettingSet ns;
  ns.id = "NS" + std::to_string(variant);
  ns.trades.push_back(std::make_shared<InterestRateSwap>("IRS1", 1e7, 0.028, schedule(0.0, 5.0, 0.5),
                                                         InterestRateSwap::Direction::PayFixed));
  ns.trades.push_back(std::make_shared<InterestRateSwap>("IRS2", 5e6, 0.031, schedule(0.25, 7.25, 1.0),
                                                         InterestRateSwap::Direction::ReceiveFixed));
  ns.trades.push_back(std::make_shared<ccr::AssetForward>("FWD", "EURUSD", 2e6, 1.12, 2.5));
  ns.trades.push_back(std::make_shared<ccr::EuropeanOption>("OPT", "SPX", ccr::OptionType::Put, -3e4, 95.0, 1.75));
  const InterestRateSwap und("UND", 8e6, 0.03, schedule(1.0, 6.0, 1.0), InterestRateSwap::Direction::PayFixed);
  const std::vector<double> ex(und.schedule().begin(), und.schedule().end() - 1);
  ns.trades.push_back(std::make_shared<ccr::BermudanSwaption>("BERM", und, ex,
                                                              variant == 2 ? ccr::Settlement::Cash : ccr::Settlement::Physical,
                                                              variant == 1 ? -1.0 : 1.0, variant == 2 ? 2 : 3));
  if (variant >= 1) {
    ccr::CollateralAgreement csa;
    csa.thresholdCounterparty = 1e5;
    csa.thresholdOwn = variant == 1 ? std::numeric_limits<double>::infinity() : 2e5;
    csa.minimumTransferAmount = 1e4;
    csa.independentAmount = variant == 2 ? -5e4 : 0.0;
    csa.marginPeriodOfRisk = 10.0 / 250.0;
    if (variant == 3) {
      csa.type = ccr::CollateralAgreement::Type::Asset;
      csa.assetIndex = 0;
      csa.haircut = 0.08;
    } else if (variant == 4) {
      csa.type = ccr::CollateralAgreement::Type::Bond;
      csa.bondMaturity = 12.0;
      csa.haircut = 0.02;
    }
    ns.collateral = csa;
  }
  return ns;
}

bool same(const std::vector<double>& a, const std::vector<double>& b) {
  if (a.size() != b.size()) return false;
  for (std::size_t i = 0; i < a.size(); ++i)
    if (!(a[i] == b[i]) && !(std::isnan(a[i]) && std::isnan(b[i]))) return false;
  return true;
}

// A trade type the C ABI does not know.
class Digital : public ccr::PathwiseTrade {
 public:
  Digital() : ccr::PathwiseTrade("DIG") {}
  double maturity() const override { return 1.0; }
  void valueAtTime(const ccr::ScenarioSet& s, std::size_t, double* out) const override {
    for (std::size_t p = 0; p < s.numPaths; ++p) out[p] = 0.0;
  }
};

}  // namespace

int main() {
  int failures = 0;
  const ccr::TimeGrid base(schedule(0.0, 8.0, 0.25));
  for (int variant = 0; variant <= 4; ++variant) {
    ccr::SimulationConfig cfg;
    cfg.numPaths = variant == 0 ? 300 : 200;
    cfg.seed = 42 + static_cast<std::uint64_t>(variant);
    cfg.antithetic = variant != 3;
    const ccr::ExposureEngine engine(generator(variant != 0), cfg, base);
    const ccr::NettingSet ns = portfolio(variant);
    const double q = variant == 2 ? 0.99 : 0.95;
    const ccr::ExposureResult ref = engine.run(ns, q);
    const ccr::ExposureResult got = ccr::hip::run(oracleEntries(), engine, ns, q);
    const ccr::ExposureProfile &a = ref.profile, &b = got.profile;
    const bool ok = same(a.times, b.times) && same(a.expectedValue, b.expectedValue) &&
                    same(a.expectedExposure, b.expectedExposure) &&
                    same(a.expectedNegativeExposure, b.expectedNegativeExposure) &&
                    same(a.potentialFutureExposure, b.potentialFutureExposure) &&
                    same(a.discountedExpectedExposure, b.discountedExpectedExposure) &&
                    same(a.discountedExpectedNegativeExposure, b.discountedExpectedNegativeExposure) &&
                    a.pfeQuantile == b.pfeQuantile && ref.reportingIndices == got.reportingIndices &&
                    same(ref.scenarios.grid.times(), got.scenarios.grid.times()) &&
                    got.scenarios.numPaths == cfg.numPaths;
    double worst = 0.0;
    for (std::size_t i = 0; i < a.expectedExposure.size() && i < b.expectedExposure.size(); ++i)
      worst = std::fmax(worst, std::fabs(a.expectedExposure[i] - b.expectedExposure[i]));
    std::printf("case %d: %zu trades%s, %zu reporting / %zu simulation dates, EE[last] %.10e: %s (max |dEE| %.3e)\n",
                variant, ns.trades.size(), ns.collateral ? " + CSA" : "", a.times.size(), ref.scenarios.grid.size(),
                a.expectedExposure.empty() ? 0.0 : a.expectedExposure.back(), ok ? "bit-identical" : "MISMATCH", worst);
    if (!ok) ++failures;
  }
  // Unsupported trade type: std::invalid_argument from the binding, before any C call.
  {
    ccr::SimulationConfig cfg;
    cfg.numPaths = 10;
    const ccr::ExposureEngine engine(generator(false), cfg, base);
    ccr::NettingSet ns;
    ns.trades.push_back(std::make_shared<Digital>());
    bool threw = false;
    try {
      (void)ccr::hip::run(oracleEntries(), engine, ns, 0.95);
    } catch (const std::invalid_argument& e) {
      threw = true;
      std::printf("unsupported trade: std::invalid_argument (%s)\n", e.what());
    }
    if (!threw) {
      std::printf("unsupported trade: NOT rejected\n");
      ++failures;
    }
  }
  if (failures == 0) std::printf("binding: all cases bit-identical\n");
  return failures == 0 ? 0 : 1;
}
