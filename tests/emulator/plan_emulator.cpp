// plan_emulator.cpp -- TEST INFRASTRUCTURE ONLY. Runs the product's plan builder (plan.cpp) and
// the product's per-path kernel body (simulatePath in path_body.hpp) on the CPU, path by path,
// in FP64, so the plan tables (merged terms, fixing slots, margin-call ring, grids) can be pinned
// against the FP64 oracle without a GPU. Per-date sums are kept in path order; the PFE uses a
// sort and the interpolation rule of ccr::quantile (include/ccr/core/math.hpp:39-47).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../zlatko_amd/csrc/path_body.hpp"
#include "../../zlatko_amd/csrc/plan.hpp"

using namespace ccrhip;

extern "C" int32_t emu_run_exposure(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg,
                                    const double* base, int32_t nbase, const ccr_netting_set* sets, int32_t nsets,
                                    double q, ccr_exposure_out* outs, char* err, int32_t cap) {
  try {
    HostPlan h = buildPlan(*gen, *cfg, base, nbase, sets, nsets, q, CCR_FP64, 256);
    std::vector<double> stream(static_cast<size_t>(h.stream_len));
    generateStream(cfg->seed, h.stream_len, stream.data());
    std::vector<double> steps = h.steps, dreal = h.date_reals, bonds = h.bonds, fixed = h.fixed;
    PathArgs<double> a{&h.g, h.sets.data(), h.block_set.data(), stream.data(), steps.data(), h.dates.data(),
                       dreal.data(), bonds.data(), fixed.data(), h.fixstore.data(), h.poly.data(), nullptr, nullptr,
                       h.max_rep, h.g.max_ring, h.price_rows, h.g.max_slots, h.g.nrates, h.g.nassets, h.g.ncredits};
    const int P = h.g.num_paths;
    std::vector<double> scratch(static_cast<size_t>(h.g.max_ring + h.price_rows + h.g.max_slots + 1));
    for (int s = 0; s < nsets; ++s) {
      const SetDev& S = h.sets[static_cast<size_t>(s)];
      const int R = S.num_rep;
      if (outs[s].capacity < R) throw Error(CCR_ERR_INVALID_ARGUMENT, "capacity");
      std::vector<double> sums(static_cast<size_t>(R) * kStats, 0.0);
      std::vector<double> rows(static_cast<size_t>(R) * P, 0.0);
      for (int p = 0; p < P; ++p) {
        std::fill(scratch.begin(), scratch.end(), 0.0);
        // The same shape specialisation the GPU launch picks.
        withShape(h.g.nrates, h.g.nassets, h.g.ncredits, [&](auto shape) {
          simulatePath<double, decltype(shape)>(a, &S, p, scratch.data(), 1, [&](int rr, const double* st, double vp) {
            for (int k = 0; k < kStats; ++k) sums[static_cast<size_t>(rr) * kStats + k] += st[k];
            rows[static_cast<size_t>(rr) * P + p] = vp;
          });
          return 0;
        });
      }
      ccr_exposure_out& o = outs[s];
      o.num_dates = R;
      for (int r = 0; r < R; ++r) {
        const double* st = &sums[static_cast<size_t>(r) * kStats];
        const double n = P;
        if (o.times) o.times[r] = h.rep_times[static_cast<size_t>(s)][static_cast<size_t>(r)];
        if (o.expected_value) o.expected_value[r] = st[0] / n;
        if (o.expected_exposure) o.expected_exposure[r] = st[1] / n;
        if (o.expected_negative_exposure) o.expected_negative_exposure[r] = st[2] / n;
        if (o.discounted_expected_exposure) o.discounted_expected_exposure[r] = st[3] / n;
        if (o.discounted_expected_negative_exposure) o.discounted_expected_negative_exposure[r] = st[4] / n;
        if (o.mean_deflator) o.mean_deflator[r] = st[5] / n;
        if (o.mean_survival) o.mean_survival[r] = st[6] / n;
        if (o.cva_increments) o.cva_increments[r] = st[7] / n;
        if (o.dva_increments) o.dva_increments[r] = st[8] / n;
        std::vector<double> v(rows.begin() + static_cast<long>(r) * P, rows.begin() + static_cast<long>(r + 1) * P);
        if (o.exposure_rows) std::copy(v.begin(), v.end(), o.exposure_rows + static_cast<size_t>(r) * P);
        std::sort(v.begin(), v.end());
        const double hq = q * static_cast<double>(P - 1);
        const size_t lo = static_cast<size_t>(std::floor(hq)), hi = std::min(lo + 1, static_cast<size_t>(P - 1));
        if (o.potential_future_exposure) o.potential_future_exposure[r] = v[lo] + (hq - static_cast<double>(lo)) * (v[hi] - v[lo]);
      }
    }
    if (err && cap > 0) err[0] = 0;
    return CCR_OK;
  } catch (const Error& e) {
    if (err && cap > 0) std::snprintf(err, static_cast<size_t>(cap), "%s", e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (err && cap > 0) std::snprintf(err, static_cast<size_t>(cap), "%s", e.what());
    return CCR_ERR_RUNTIME;
  }
}

// Plan dimensions (no execution): max_ring, max_slots, max_dates, max_rep, total_terms, stream_len.
extern "C" int32_t emu_plan_stats(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg,
                                  const double* base, int32_t nbase, const ccr_netting_set* sets, int32_t nsets,
                                  int64_t* out6) {
  try {
    HostPlan h = buildPlan(*gen, *cfg, base, nbase, sets, nsets, 0.95, CCR_FP64, 256);
    out6[0] = h.g.max_ring;
    out6[1] = h.g.max_slots;
    out6[2] = h.max_dates;
    out6[3] = h.max_rep;
    out6[4] = h.total_terms;
    out6[5] = h.stream_len;
    return CCR_OK;
  } catch (const std::exception&) {
    return CCR_ERR_RUNTIME;
  }
}

// The product's FP64 bond-term exponential, evaluated on the host.
extern "C" void emu_exp_bond(const double* x, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = expBond(x[i]);
}
