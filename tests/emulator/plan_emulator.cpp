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

namespace {

// Masked polynomial regression with the moments summed in path order (the order of
// PolynomialRegression::fit, src/core/regression.cpp:37-80), through the product's RegParams /
// regSolve arithmetic. Returns the record center, scale, c0..c3, degree.
void fitRecord(const std::vector<double>& x, const std::vector<double>& y, const std::vector<char>& mask, int degree,
               double* rec) {
  double n = 0, sx = 0, sxx = 0, sy = 0;
  for (size_t i = 0; i < x.size(); ++i) {
    if (!mask[i]) continue;
    n += 1.0;
    sx += x[i];
    sxx += x[i] * x[i];
    sy += y[i];
  }
  const RegParams r = regParams(n, sx, sxx, degree);
  double gram[10] = {0}, xty[4] = {0};
  if (r.deg > 0)
    for (size_t i = 0; i < x.size(); ++i) {
      if (!mask[i]) continue;
      double phi[4];
      regRow((x[i] - r.center) / r.scale, phi);
      for (int a = 0; a <= r.deg; ++a) {
        xty[a] += phi[a] * y[i];
        for (int b = 0; b <= a; ++b) gram[triIndex(a, b)] += phi[a] * phi[b];
      }
    }
  double coef[4];
  regSolve(r, gram, xty, sy, coef);
  rec[0] = r.n == 0.0 ? 0.0 : r.center;
  rec[1] = r.scale;
  for (int i = 0; i < 4; ++i) rec[2 + i] = coef[i];
  rec[6] = static_cast<double>(r.deg);
  rec[7] = 0.0;
}

// Regression basis of the plans built by the emulator (emu_set_regression_basis).
int gBasis = -1;

// The regression engine on the CPU: every record of the plan (exercise records per swaption from
// the last exercise down, then the continuation records) fitted with the product's
// basisParams / basisRow / basisSolve arithmetic, moments and Gram summed in path order.
struct EngineState {
  std::vector<double> rfit;
};

void runEngine(HostPlan& h, PathArgs<double>& a, const std::vector<std::vector<std::vector<double>>>& ev,
               std::vector<double>& cont, std::vector<int16_t>& exidx, EngineState& eng) {
  const int P = h.g.num_paths;
  const double* X = a.states;
  const double* Dv = a.states + h.state_len;
  eng.rfit.assign(h.rrecs.size() * kBasisStride, 0.0);
  std::vector<std::vector<std::vector<double>>> strategy(h.berms.size());
  for (size_t b = 0; b < h.berms.size(); ++b)
    strategy[b].assign(static_cast<size_t>(h.berms[b].ne), std::vector<double>(static_cast<size_t>(P), 0.0));
  auto features = [&](const RegRec& rr, double x, double& u1, double& u2) {
    u1 = x;
    u2 = 0.0;
    if (h.basis != kBasisSwapRateAnnuity) return;
    const SetDev& S = h.sets[static_cast<size_t>(h.berms[static_cast<size_t>(rr.b)].set)];
    auto noFix = [](int) { return 0.0; };
    const double* bb = a.bonds + 2 * static_cast<int64_t>(S.bond_off);
    const double* ff = a.fixed + 4 * static_cast<int64_t>(S.fixed_off);
    const double* pp = a.poly + S.poly_off;
    const double A = evalRange<double>(h.saterms[static_cast<size_t>(rr.ann)], x, bb, ff, pp, noFix);
    const double F = evalRange<double>(h.saterms[static_cast<size_t>(rr.flt)], x, bb, ff, pp, noFix);
    u1 = F / A;
    u2 = A;
  };
  // Sample of record r on path p: in / features / target numerator.
  auto sample = [&](int r, int p, bool& in, double& u1, double& u2, double& y, double& e, double& D, double& num) {
    const RegRec& rr = h.rrecs[static_cast<size_t>(r)];
    const BermDev& bm = h.berms[static_cast<size_t>(rr.b)];
    const SetDev& S = h.sets[static_cast<size_t>(bm.set)];
    const int64_t at = S.state_off + static_cast<int64_t>(rr.j) * P + p;
    D = Dv[at];
    const auto& st = strategy[static_cast<size_t>(rr.b)];
    if (rr.kind == 0) {
      e = ev[static_cast<size_t>(rr.b)][static_cast<size_t>(rr.k)][static_cast<size_t>(p)];
      in = e > 0.0;
      num = rr.k + 1 < bm.ne ? st[static_cast<size_t>(rr.k + 1)][static_cast<size_t>(p)] : 0.0;
    } else {
      e = 0.0;
      in = true;
      num = st[static_cast<size_t>(rr.k)][static_cast<size_t>(p)];
    }
    y = num / D;
    features(rr, X[at], u1, u2);
  };
  auto fit = [&](int r) {
    const RegRec& rr = h.rrecs[static_cast<size_t>(r)];
    double m[kBasisMom] = {0, 0, 0, 0, 0, 0};
    bool in;
    double u1, u2, y, e, D, num;
    for (int p = 0; p < P; ++p) {
      sample(r, p, in, u1, u2, y, e, D, num);
      if (!in) continue;
      m[0] += 1.0;
      m[1] += u1;
      m[2] += u1 * u1;
      m[3] += u2;
      m[4] += u2 * u2;
      m[5] += y;
    }
    double* rp = eng.rfit.data() + static_cast<size_t>(r) * kBasisStride;
    basisParams(h.basis, h.berms[static_cast<size_t>(rr.b)].degree, m, rp);
    if (rp[5] < 1.0) return;
    double g[kBasisGram] = {0};
    for (int p = 0; p < P; ++p) {
      sample(r, p, in, u1, u2, y, e, D, num);
      if (!in) continue;
      double phi[kBasisMax + 1];
      basisRow(h.basis, rp, u1, u2, phi);
      phi[kBasisMax] = y;
      for (int i = 0; i <= kBasisMax; ++i)
        for (int c = i; c <= kBasisMax; ++c) g[gramIndex(i, c)] += phi[i] * phi[c];
    }
    basisSolve(h.basis, g, rp);
  };
  for (int k = h.max_exercises - 1; k >= 0; --k)
    for (size_t b = 0; b < h.berms.size(); ++b) {
      const int r = h.rex[b * static_cast<size_t>(h.max_exercises) + static_cast<size_t>(k)];
      if (r < 0) continue;
      fit(r);
      const double* rp = eng.rfit.data() + static_cast<size_t>(r) * kBasisStride;
      int16_t* ex = exidx.data() + b * static_cast<size_t>(P);
      for (int p = 0; p < P; ++p) {
        bool in;
        double u1, u2, y, e, D, num;
        sample(r, p, in, u1, u2, y, e, D, num);
        double realised = num;
        if (e > 0.0 && e > basisPredict(h.basis, rp, u1, u2)) {
          ex[p] = static_cast<int16_t>(k);
          realised = D * e;
        }
        strategy[b][static_cast<size_t>(k)][static_cast<size_t>(p)] = realised;
      }
    }
  for (size_t r = static_cast<size_t>(h.nex_recs); r < h.rrecs.size(); ++r) fit(static_cast<int>(r));
  for (size_t r = 0; r < h.bdates.size(); ++r) {
    double* out = cont.data() + r * kContStride;
    for (int i = 0; i < kContStride; ++i) out[i] = 0.0;
    out[1] = 1.0;
    out[6] = -1.0;
    const int rec = h.rcont[r];
    if (rec < 0 || h.basis != kBasisState) continue;
    const double* rp = eng.rfit.data() + static_cast<size_t>(rec) * kBasisStride;
    out[0] = rp[1];
    out[1] = rp[2];
    for (int i = 0; i < 4; ++i) out[2 + i] = i <= static_cast<int>(rp[5]) ? rp[6 + i] : 0.0;
    out[6] = rp[5];
  }
}

// The American Monte Carlo stage on the CPU: states pass, exercise values, backward induction
// and continuation fits, filling cont and exidx for the valuation pass.
void runAmc(HostPlan& h, PathArgs<double>& a, std::vector<double>& states, std::vector<double>& cont,
            std::vector<int16_t>& exidx, std::vector<double>& bermv, std::vector<double>& scratch) {
  const int P = h.g.num_paths;
  states.assign(static_cast<size_t>(2 * h.state_len), 0.0);
  cont.assign(static_cast<size_t>(h.cont_len), 0.0);
  exidx.assign(h.berms.size() * static_cast<size_t>(P), 0);
  PathArgs<double> s = a;
  s.states = states.data();
  for (size_t si = 0; si < h.sets.size(); ++si) {
    const SetDev& S = h.sets[si];
    if (S.state_off < 0) continue;
    for (int p = 0; p < P; ++p) simulateStates<double>(s, &S, p, true);
  }
  const double* X = states.data();
  const double* Dv = states.data() + h.state_len;
  EngineState eng;
  std::vector<std::vector<std::vector<double>>> evAll(h.basis >= 0 ? h.berms.size() : 0);
  for (size_t b = 0; b < h.berms.size(); ++b) {
    const BermDev& bm = h.berms[b];
    const SetDev& S = h.sets[static_cast<size_t>(bm.set)];
    const int ne = bm.ne;
    std::vector<std::vector<double>> ev(static_cast<size_t>(ne), std::vector<double>(static_cast<size_t>(P)));
    for (int k = 0; k < ne; ++k) {
      const int j = h.eidx[static_cast<size_t>(bm.eidx_off + k)];
      for (int p = 0; p < P; ++p)
        ev[static_cast<size_t>(k)][static_cast<size_t>(p)] =
            evalRange<double>(h.evterms[static_cast<size_t>(bm.ev_rec + k)], X[S.state_off + static_cast<int64_t>(j) * P + p],
                              a.bonds + 2 * static_cast<int64_t>(S.bond_off), a.fixed + 4 * static_cast<int64_t>(S.fixed_off),
                              a.poly + S.poly_off, [](int) { return 0.0; });
    }
    std::fill(exidx.data() + b * static_cast<size_t>(P), exidx.data() + (b + 1) * static_cast<size_t>(P),
              static_cast<int16_t>(ne));
    if (h.basis >= 0) {
      evAll[b] = ev;
      continue;
    }
    std::vector<std::vector<double>> strategy(static_cast<size_t>(ne) + 1, std::vector<double>(static_cast<size_t>(P), 0.0));
    std::vector<double> realised(static_cast<size_t>(P), 0.0), x(static_cast<size_t>(P)), y(static_cast<size_t>(P));
    std::vector<char> itm(static_cast<size_t>(P)), all(static_cast<size_t>(P), 1);
    int16_t* ex = exidx.data() + b * static_cast<size_t>(P);
    std::fill(ex, ex + P, static_cast<int16_t>(ne));
    for (int k = ne - 1; k >= 0; --k) {
      const int j = h.eidx[static_cast<size_t>(bm.eidx_off + k)];
      const int64_t at = S.state_off + static_cast<int64_t>(j) * P;
      for (int p = 0; p < P; ++p) {
        x[static_cast<size_t>(p)] = X[at + p];
        y[static_cast<size_t>(p)] = realised[static_cast<size_t>(p)] / Dv[at + p];
        itm[static_cast<size_t>(p)] = ev[static_cast<size_t>(k)][static_cast<size_t>(p)] > 0.0;
      }
      double rec[8];
      fitRecord(x, y, itm, bm.degree, rec);
      for (int p = 0; p < P; ++p) {
        const double e = ev[static_cast<size_t>(k)][static_cast<size_t>(p)];
        if (e > 0.0 && e > regPredict(x[static_cast<size_t>(p)], rec[0], rec[1], static_cast<int>(rec[6]), rec + 2)) {
          ex[p] = static_cast<int16_t>(k);
          realised[static_cast<size_t>(p)] = Dv[at + p] * e;
        }
      }
      strategy[static_cast<size_t>(k)] = realised;
    }
    for (int j = 0; j < S.num_dates; ++j) {
      const BermDate& bd = h.bdates[static_cast<size_t>(bm.date_rec + j)];
      double* rec = cont.data() + static_cast<size_t>(bm.date_rec + j) * kContStride;
      rec[1] = 1.0;
      rec[6] = -1.0;
      if (bd.next_k >= ne || h.bdate_cd[static_cast<size_t>(bm.date_rec + j)] < 0) continue;
      const int64_t at = S.state_off + static_cast<int64_t>(j) * P;
      for (int p = 0; p < P; ++p) {
        x[static_cast<size_t>(p)] = X[at + p];
        y[static_cast<size_t>(p)] = strategy[static_cast<size_t>(bd.next_k)][static_cast<size_t>(p)] / Dv[at + p];
      }
      fitRecord(x, y, all, bm.degree, rec);
    }
  }
  if (h.basis >= 0) {
    PathArgs<double> e = a;
    e.states = states.data();
    runEngine(h, e, evAll, cont, exidx, eng);
  }
  BasisView view;
  view.basis = h.basis;
  if (h.basis == kBasisSwapRateAnnuity) {
    view.rrecs = h.rrecs.data();
    view.rcont = h.rcont.data();
    view.rfit = eng.rfit.data();
    view.saterms = h.saterms.data();
  }
  const int ncd = static_cast<int>(h.cdates.size() / 2);
  bermv.assign(static_cast<size_t>(ncd) * P, 0.0);
  for (int d = 0; d < ncd; ++d) {
    const SetDev& S = h.sets[static_cast<size_t>(h.cdates[2 * static_cast<size_t>(d)])];
    const int j = h.cdates[2 * static_cast<size_t>(d) + 1];
    for (int p = 0; p < P; ++p) {
      const double* bb = a.bonds + 2 * static_cast<int64_t>(S.bond_off);
      const double* ff = a.fixed + 4 * static_cast<int64_t>(S.fixed_off);
      const int lb = h.clive_rng[2 * static_cast<size_t>(d)], le = h.clive_rng[2 * static_cast<size_t>(d) + 1];
      if (h.basis == kBasisSwapRateAnnuity) {
        bermv[static_cast<size_t>(d) * P + p] = bermudanSum<double, true>(
            S, j, p, P, cont.data(), exidx.data(), states.data(), bb, ff, a.poly + S.poly_off, h.clive.data(), lb, le,
            view);
      } else {
        // The state basis as the GPU values it (bermudanSumN; one path per call on the host).
        const int64_t one[1] = {p};
        double v[1];
        bermudanSumN<double, 1>(S, j, one, P, cont.data(), exidx.data(), states.data(), bb, ff, a.poly + S.poly_off,
                                h.clive.data(), lb, le, v);
        bermv[static_cast<size_t>(d) * P + p] = v[0];
      }
    }
  }
  a.bermv = bermv.data();
}

int32_t runExposure(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg, const double* base,
                    int32_t nbase, const ccr_netting_set* sets, int32_t nsets, double q, ccr_exposure_out* outs,
                    char* err, int32_t cap, bool book);

}  // namespace

// Bermudan regressions of later runs: -1 the shared-design state basis, 0 / 1 the regression
// engine with the state / swap-rate-annuity basis.
extern "C" void emu_set_regression_basis(int32_t basis) { gBasis = basis; }

extern "C" int32_t emu_run_exposure(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg,
                                    const double* base, int32_t nbase, const ccr_netting_set* sets, int32_t nsets,
                                    double q, ccr_exposure_out* outs, char* err, int32_t cap) {
  return runExposure(gen, cfg, base, nbase, sets, nsets, q, outs, err, cap, false);
}

// The same with the plan built in book mode (common grids for every set).
extern "C" int32_t emu_run_book(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg,
                                const double* base, int32_t nbase, const ccr_netting_set* sets, int32_t nsets,
                                double q, ccr_exposure_out* outs, char* err, int32_t cap) {
  return runExposure(gen, cfg, base, nbase, sets, nsets, q, outs, err, cap, true);
}

namespace {

int32_t runExposure(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg, const double* base,
                    int32_t nbase, const ccr_netting_set* sets, int32_t nsets, double q, ccr_exposure_out* outs,
                    char* err, int32_t cap, bool book) {
  try {
    HostPlan h = buildPlan(*gen, *cfg, base, nbase, sets, nsets, q, CCR_FP64, 256, false, book, false, nullptr,
                           nullptr, gBasis);
    std::vector<double> stream(static_cast<size_t>(h.stream_len));
    generateStream(cfg->seed, h.stream_len, stream.data());
    std::vector<double> steps = h.steps, dreal = h.date_reals, bonds = h.bonds, fixed = h.fixed;
    PathArgs<double> a{};
    a.g = &h.g;
    a.sets = h.sets.data();
    a.block_set = h.block_set.data();
    a.stream = stream.data();
    a.steps = steps.data();
    a.dates = h.dates.data();
    a.date_reals = dreal.data();
    a.bonds = bonds.data();
    a.fixed = fixed.data();
    a.fixstore = h.fixstore.data();
    a.poly = h.poly.data();
    a.opts = h.opts.data();
    a.max_rep = h.max_rep;
    a.ring = h.g.max_ring;
    a.price_rows = h.price_rows;
    a.slots = h.g.max_slots;
    a.nrates = h.g.nrates;
    a.nassets = h.g.nassets;
    a.ncredits = h.g.ncredits;
    a.state_len = h.state_len;
    const int P = h.g.num_paths;
    std::vector<double> scratch(static_cast<size_t>(h.g.max_ring + h.price_rows + h.g.max_slots + 1));
    std::vector<double> states, cont, bermv;
    std::vector<int16_t> exidx;
    if (!h.berms.empty()) runAmc(h, a, states, cont, exidx, bermv, scratch);
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

}  // namespace

// Per-trade plan of one netting set, run through the kernel body's per-trade path (the generic
// shape with per-trade values, as the GPU launch picks): stats[(i * kTradeStats + m) * capacity + r]
// are the per-trade statistics summed in path order and divided by the paths.
extern "C" int32_t emu_run_trade_statistics(const ccr_scenario_generator* gen, const ccr_simulation_config* cfg,
                                            const double* base, int32_t nbase, const ccr_netting_set* set,
                                            double* stats, int32_t capacity, char* err, int32_t cap) {
  try {
    HostPlan h = buildPlan(*gen, *cfg, base, nbase, set, 1, 0.95, CCR_FP64, 256, true);
    std::vector<double> stream(static_cast<size_t>(h.stream_len));
    generateStream(cfg->seed, h.stream_len, stream.data());
    PathArgs<double> a{};
    a.g = &h.g;
    a.sets = h.sets.data();
    a.block_set = h.block_set.data();
    a.stream = stream.data();
    a.steps = h.steps.data();
    a.dates = h.dates.data();
    a.date_reals = h.date_reals.data();
    a.bonds = h.bonds.data();
    a.fixed = h.fixed.data();
    a.fixstore = h.fixstore.data();
    a.poly = h.poly.data();
    a.opts = h.opts.data();
    a.max_rep = h.max_rep;
    a.ring = h.g.max_ring;
    a.price_rows = h.price_rows;
    a.slots = h.g.max_slots;
    a.nrates = h.g.nrates;
    a.nassets = h.g.nassets;
    a.ncredits = h.g.ncredits;
    a.tstride = kTradeStats * h.max_groups;
    a.tdates = h.tdates.data();
    a.tdreals = h.tdreals.data();
    const SetDev& S = h.sets[0];
    const int P = h.g.num_paths, R = S.num_rep, G = S.num_groups;
    if (capacity < R) throw Error(CCR_ERR_INVALID_ARGUMENT, "capacity");
    std::vector<double> sums(static_cast<size_t>(G) * kTradeStats * R, 0.0);
    std::vector<double> scratch(static_cast<size_t>(h.g.max_ring + h.price_rows + h.g.max_slots + G + 1));
    for (int p = 0; p < P; ++p) {
      std::fill(scratch.begin(), scratch.end(), 0.0);
      simulatePath<double, GenericShape, true, true>(
          a, &S, p, scratch.data(), 1, [](int, const double*, double) {}, [](int, double, double) {},
          [&](int rr, double V, double D, const double* tv, int stride) {
            const double on = V > 0.0 ? 1.0 : 0.0, vp = V > 0.0 ? V : 0.0;
            for (int i = 0; i < G; ++i) {
              const double vi = tv[i * stride], rest = V - vi, vip = vi > 0.0 ? vi : 0.0;
              const double v5[kTradeStats] = {D * vi * on, vi * on, D * (vp - (rest > 0.0 ? rest : 0.0)), D * vip, vip};
              for (int m = 0; m < kTradeStats; ++m)
                sums[(static_cast<size_t>(i) * kTradeStats + static_cast<size_t>(m)) * R + static_cast<size_t>(rr)] += v5[m];
            }
          });
    }
    for (int i = 0; i < G; ++i)
      for (int m = 0; m < kTradeStats; ++m)
        for (int r = 0; r < R; ++r)
          stats[(static_cast<size_t>(i) * kTradeStats + static_cast<size_t>(m)) * capacity + static_cast<size_t>(r)] =
              sums[(static_cast<size_t>(i) * kTradeStats + static_cast<size_t>(m)) * R + static_cast<size_t>(r)] / P;
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
