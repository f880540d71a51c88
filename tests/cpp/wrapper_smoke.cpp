// Smoke program of the C++ wrapper (include/ccr_hip.hpp): one netting set of two swaps and an FX
// forward through ccr::hip::ExposureEngine::run; prints the reporting dates, EE and the EPE, or the
// wrapper's exception message.
#include <cstdio>

#include "ccr_hip.hpp"

int main() {
  double ct[3] = {1.0, 3.0, 10.0}, cz[3] = {0.02, 0.025, 0.03}, fz[1] = {0.01}, ft[1] = {1.0};
  ccr_lognormal_asset fx{};
  fx.spot = 1.1;
  fx.volatility = 0.12;
  fx.carry.times = ft;
  fx.carry.zero_rates = fz;
  fx.carry.size = 1;
  fx.foreign_rate = -1;
  double corr[4] = {1.0, 0.3, 0.3, 1.0};
  ccr_scenario_generator g{};
  g.rates.curve.times = ct;
  g.rates.curve.zero_rates = cz;
  g.rates.curve.size = 3;
  g.rates.mean_reversion = 0.05;
  g.rates.volatility = 0.01;
  g.num_assets = 1;
  g.assets = &fx;
  g.correlation = corr;
  const ccr_simulation_config c{2000, 42, 1};
  double s1[5] = {0.0, 0.5, 1.0, 1.5, 2.0}, s2[7] = {0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0};
  ccr_trade tr[3] = {};
  tr[0].kind = CCR_TRADE_SWAP;
  tr[0].pay_fixed = 1;
  tr[0].notional = 1e6;
  tr[0].fixed_rate = 0.025;
  tr[0].schedule = s1;
  tr[0].schedule_size = 5;
  tr[1] = tr[0];
  tr[1].pay_fixed = 0;
  tr[1].notional = 5e5;
  tr[1].fixed_rate = 0.03;
  tr[1].schedule = s2;
  tr[1].schedule_size = 7;
  tr[2].kind = CCR_TRADE_ASSET_FORWARD;
  tr[2].notional = 4e5;
  tr[2].fixed_rate = 1.1;
  tr[2].maturity = 2.5;
  ccr_netting_set ns{};
  ns.num_trades = 3;
  ns.trades = tr;
  ns.counterparty_credit = -1;
  ns.own_credit = -1;
  try {
    auto ctx = std::make_shared<ccr::hip::Context>();
    std::vector<double> base;
    for (int i = 0; i <= 12; ++i) base.push_back(0.25 * i);
    ccr::hip::ExposureEngine engine(ctx, g, c, base);
    const ccr::hip::ExposureResult r = engine.run(ns, 0.95);
    for (size_t j = 0; j < r.reportingTimes.size(); ++j)
      std::printf("%.17g %.17g\n", r.reportingTimes[j], r.profile.expectedExposure[j]);
    std::printf("EPE %.17g\n", r.profile.expectedPositiveExposure(3.0));
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
  }
  return 0;
}
