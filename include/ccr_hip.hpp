// ccr_hip.hpp -- header-only C++17 wrapper of the C ABI (ccr_hip.h) with the shape of the
// reference's engine interface:
//   ccr::ExposureEngine(generator, config, baseGrid); ExposureResult run(const NettingSet&,
//   double pfeQuantile = 0.95) const; reportingGrid / simulationGrid
//   (include/ccr/exposure/exposure_engine.hpp:48-66)
// and the reference's exception types for failures (std::invalid_argument, std::out_of_range,
// std::runtime_error; exposure_engine.cpp:51,66). Inputs are the C structs of ccr_hip.h, which a
// caller fills from its ccr:: objects (see the integration notes for the field mapping).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ccr_hip.h"

namespace ccr {
namespace hip {

// Throws the exception type the reference raises for the same condition.
inline void throwOnError(int32_t code, const char* message) {
  if (code == CCR_OK) return;
  const std::string m = message ? message : "ccr_hip error";
  if (code == CCR_ERR_INVALID_ARGUMENT) throw std::invalid_argument(m);
  if (code == CCR_ERR_OUT_OF_RANGE) throw std::out_of_range(m);
  throw std::runtime_error(m);
}

// ccr::ExposureProfile (include/ccr/exposure/exposure_profile.hpp) on the reporting dates, plus
// the pathwise CVA / DVA increments of the fused backend.
struct ExposureProfile {
  std::vector<double> times, expectedValue, expectedExposure, expectedNegativeExposure, potentialFutureExposure,
      discountedExpectedExposure, discountedExpectedNegativeExposure, meanDeflator, meanSurvival, cvaIncrements,
      dvaIncrements;
  double expectedPositiveExposure(double horizon) const {
    return ccr_expected_positive_exposure(times.data(), expectedExposure.data(), static_cast<int32_t>(times.size()),
                                          horizon);
  }
  std::vector<double> effectiveExpectedExposure() const {
    std::vector<double> eff(expectedExposure.size());
    ccr_effective_expected_exposure(expectedExposure.data(), static_cast<int32_t>(eff.size()), eff.data());
    return eff;
  }
  double effectiveExpectedPositiveExposure(double horizon) const {
    return ccr_effective_expected_positive_exposure(times.data(), expectedExposure.data(),
                                                    static_cast<int32_t>(times.size()), horizon);
  }
};

// ccr::ExposureResult (exposure_engine.hpp:26-37): the profile on the reporting grid.
struct ExposureResult {
  std::vector<double> reportingTimes;
  ExposureProfile profile;
  double pathwiseCva = 0.0, pathwiseDva = 0.0;
};

// One device context (stream, events) -- single-threaded, as the C ABI.
class Context {
 public:
  explicit Context(int device = 0, int precision = CCR_FP64, int blockSize = 0) {
    const ccr_hip_options o{device, precision, blockSize};
    ccr_hip_ctx* c = nullptr;
    const int32_t rc = ccr_hip_create(&o, &c);
    throwOnError(rc, ccr_hip_last_error(nullptr));
    ctx_.reset(c);
  }
  ccr_hip_ctx* get() const { return ctx_.get(); }
  void check(int32_t code) const { throwOnError(code, ccr_hip_last_error(ctx_.get())); }

 private:
  struct Del {
    void operator()(ccr_hip_ctx* c) const { ccr_hip_destroy(c); }
  };
  std::unique_ptr<ccr_hip_ctx, Del> ctx_;
};

class ExposureEngine {
 public:
  ExposureEngine(std::shared_ptr<Context> context, const ccr_scenario_generator& generator,
                 const ccr_simulation_config& config, std::vector<double> baseGrid, int precision = CCR_FP64)
      : ctx_(std::move(context)), gen_(generator), cfg_(config), base_(std::move(baseGrid)), precision_(precision) {}

  std::vector<double> reportingGrid(const ccr_netting_set& set) const { return grid(set, false); }
  std::vector<double> simulationGrid(const ccr_netting_set& set) const { return grid(set, true); }

  // ExposureEngine::run (exposure_engine.cpp:50-81) for one netting set.
  ExposureResult run(const ccr_netting_set& set, double pfeQuantile = 0.95) const {
    return runAll(&set, 1, pfeQuantile).front();
  }
  // Several netting sets in one launch (one plan; each set on its own grid and stream slice).
  std::vector<ExposureResult> runAll(const ccr_netting_set* sets, int32_t count, double pfeQuantile = 0.95) const {
    std::vector<ExposureResult> res(static_cast<size_t>(count));
    std::vector<ccr_exposure_out> outs(static_cast<size_t>(count));
    for (int32_t s = 0; s < count; ++s) {
      const auto times = reportingGrid(sets[s]);
      ExposureProfile& p = res[static_cast<size_t>(s)].profile;
      for (auto* v : {&p.times, &p.expectedValue, &p.expectedExposure, &p.expectedNegativeExposure,
                      &p.potentialFutureExposure, &p.discountedExpectedExposure, &p.discountedExpectedNegativeExposure,
                      &p.meanDeflator, &p.meanSurvival, &p.cvaIncrements, &p.dvaIncrements})
        v->assign(times.size(), 0.0);
      ccr_exposure_out& o = outs[static_cast<size_t>(s)];
      o = ccr_exposure_out{};
      o.capacity = static_cast<int32_t>(times.size());
      o.times = p.times.data();
      o.expected_value = p.expectedValue.data();
      o.expected_exposure = p.expectedExposure.data();
      o.expected_negative_exposure = p.expectedNegativeExposure.data();
      o.potential_future_exposure = p.potentialFutureExposure.data();
      o.discounted_expected_exposure = p.discountedExpectedExposure.data();
      o.discounted_expected_negative_exposure = p.discountedExpectedNegativeExposure.data();
      o.mean_deflator = p.meanDeflator.data();
      o.mean_survival = p.meanSurvival.data();
      o.cva_increments = p.cvaIncrements.data();
      o.dva_increments = p.dvaIncrements.data();
    }
    ctx_->check(ccr_hip_run_exposure(ctx_->get(), &gen_, &cfg_, base_.data(), static_cast<int32_t>(base_.size()), sets,
                                     count, pfeQuantile, precision_, outs.data()));
    for (auto& r : res) {
      r.reportingTimes = r.profile.times;
      for (size_t j = 1; j < r.profile.cvaIncrements.size(); ++j) {
        r.pathwiseCva += r.profile.cvaIncrements[j];
        r.pathwiseDva += r.profile.dvaIncrements[j];
      }
    }
    return res;
  }

 private:
  std::vector<double> grid(const ccr_netting_set& set, bool simulation) const {
    int32_t n = 0;
    auto f = simulation ? ccr_hip_simulation_grid : ccr_hip_reporting_grid;
    ctx_->check(f(base_.data(), static_cast<int32_t>(base_.size()), &set, nullptr, 0, &n));
    std::vector<double> t(static_cast<size_t>(n));
    ctx_->check(f(base_.data(), static_cast<int32_t>(base_.size()), &set, t.data(), n, &n));
    return t;
  }

  std::shared_ptr<Context> ctx_;
  ccr_scenario_generator gen_;
  ccr_simulation_config cfg_;
  std::vector<double> base_;
  int precision_;
};

}  // namespace hip
}  // namespace ccr
