// plan.cpp -- host plan builder (FP64). See plan.hpp and layout.hpp.
//
// Valuation terms. On date t every domestic or foreign swap and every forward is a sum of
// zero-coupon bonds P_r(t, T | x_r) = A_r(t, T) exp(-B_r(t, T) x_r) of its rate r, times the FX
// spot of r for foreign rates (interest_rate_swap.cpp:47-78, asset_forward.cpp:13-25):
//   unfixed period (ts >= t):   s N (P(t,ts) - P(t,te) - K tau P(t,te))
//   fixed period  (ts <  t):    s N (P(tk,ts|x_k) / P(tk,te|x_k) - 1 - K tau) P(t,te)
//                             = s N R A(t,te) exp(-B(t,te) x_t - dB x_k) - s N (1 + K tau) P(t,te)
//   forward:                    N (S Q(t,T) - K P(t,T))
// with k = indexAtOrBefore(ts), R = A(tk,ts)/A(tk,te) and dB = B(tk,ts) - B(tk,te) (A = 1, B = 0
// where the reference's zeroBond returns 1 for T <= t). Terms of one date are merged across the
// netting set per (rate, maturity) and per (rate, fixing date, ts, te), so the kernel evaluates one
// exponential per distinct cash-flow date instead of 2-3 zero bonds per period per trade. The
// x_k a fixed coupon needs is kept per path in a slot recorded on the fixing date; slots are
// reused once their last coupon is paid (interval colouring).
#include "plan.hpp"

#include <cstring>
#include <map>
#include <random>
#include <tuple>

namespace ccrhip {

namespace {

struct Trade {
  int kind = CCR_TRADE_SWAP;
  int rate = 0;  // rate group of the swap (0 domestic, f + 1 foreign)
  double sign = 1.0, notional = 0.0, K = 0.0, maturity = 0.0;
  std::vector<double> sched, fsched;
  int asset = 0;
  // Bermudan swaption (the swap fields hold the underlying).
  std::vector<double> exercise;
  bool cash = false;
  double position = 1.0;
  int degree = 3;
  int optionType = CCR_OPTION_CALL;  // European option (asset, notional, K = strike, maturity = expiry)
};

void checkSchedule(const std::vector<double>& s) {
  if (s.size() < 2) invalid("InterestRateSwap: schedule needs at least two dates");
  if (s.front() < 0.0) invalid("InterestRateSwap: schedule must start at t >= 0");
  for (std::size_t i = 1; i < s.size(); ++i)
    if (s[i] <= s[i - 1]) invalid("InterestRateSwap: schedule must be strictly increasing");
}

std::vector<double> eventTimes(const ccr_trade& d) {
  std::vector<double> e;
  if (d.kind == CCR_TRADE_ASSET_FORWARD || d.kind == CCR_TRADE_EUROPEAN_OPTION) return {d.maturity};
  if (d.schedule && d.schedule_size > 0) e.assign(d.schedule, d.schedule + d.schedule_size);
  if (d.float_schedule && d.float_schedule_size > 0)
    e.insert(e.end(), d.float_schedule, d.float_schedule + d.float_schedule_size);
  if (d.kind == CCR_TRADE_BERMUDAN_SWAPTION && d.exercise_times && d.num_exercises > 0)
    e.insert(e.end(), d.exercise_times, d.exercise_times + d.num_exercises);
  return e;
}

struct Models {
  std::vector<HullWhite> rates;  // [0] domestic, [f + 1] foreign f
  std::vector<Curve> carry;      // deterministic carry per asset
  std::vector<int> assetRate;    // 0 = deterministic carry, f + 1 = foreign rate f
  std::vector<double> vol, spot;
  std::vector<CirModel> credits;
  std::vector<int> fxOf;         // per rate group: converting asset (-1 none; [0] unused)
  std::vector<double> quanto;    // per rate group: drift of x under the domestic measure
  std::vector<double> chol;
  int ncorr = 0, nz = 0, foreignOff = 0;
  std::vector<int> creditOff;
};

Models makeModels(const ccr_scenario_generator& g) {
  Models m;
  if (g.num_foreign_rates < 0 || g.num_assets < 0 || g.num_credits < 0) invalid("ScenarioGenerator: negative count");
  if (g.num_foreign_rates + 1 > kMaxRates || g.num_assets > kMaxAssets || g.num_credits > kMaxCredits)
    fail(CCR_ERR_UNSUPPORTED, "ScenarioGenerator: at most 3 foreign rates, 4 assets and 2 credits");
  m.rates.emplace_back(g.rates);
  for (int f = 0; f < g.num_foreign_rates; ++f) m.rates.emplace_back(g.foreign_rates[f]);
  const int F = g.num_foreign_rates;
  for (int a = 0; a < g.num_assets; ++a) {
    const ccr_lognormal_asset& as = g.assets[a];
    if (as.spot <= 0.0) invalid("LognormalAsset: spot must be positive");
    if (as.volatility < 0.0) invalid("LognormalAsset: volatility must be non-negative");
    if (as.foreign_rate >= F) invalid("LognormalAsset: unknown foreign rate");
    m.spot.push_back(as.spot);
    m.vol.push_back(as.volatility);
    m.assetRate.push_back(as.foreign_rate < 0 ? 0 : as.foreign_rate + 1);
    m.carry.push_back(as.foreign_rate < 0 ? Curve(as.carry) : Curve());
  }
  for (int c = 0; c < g.num_credits; ++c) {
    m.credits.emplace_back(g.credits[c]);
    if (m.credits.back().sub > kMaxSub) fail(CCR_ERR_UNSUPPORTED, "CirIntensity: at most 8 substeps");
  }
  const int n = 1 + F + g.num_assets + g.num_credits;
  m.ncorr = n;
  std::vector<double> corr(static_cast<std::size_t>(n * n), 0.0);
  if (g.correlation) {
    for (int i = 0; i < n * n; ++i) corr[static_cast<std::size_t>(i)] = g.correlation[i];
  } else {
    for (int i = 0; i < n; ++i) corr[static_cast<std::size_t>(i * n + i)] = 1.0;
  }
  for (int i = 0; i < n; ++i)
    if (std::fabs(corr[static_cast<std::size_t>(i * n + i)] - 1.0) > 1e-12)
      invalid("ScenarioGenerator: correlation matrix must have unit diagonal");
  m.chol = cholesky(corr, static_cast<std::size_t>(n));
  m.nz = n + 1;
  for (int c = 0; c < g.num_credits; ++c) {
    m.creditOff.push_back(m.nz);
    m.nz += m.credits[static_cast<std::size_t>(c)].sub - 1;
  }
  m.foreignOff = m.nz;
  m.nz += F;
  m.fxOf.assign(static_cast<std::size_t>(F + 1), -1);
  m.quanto.assign(static_cast<std::size_t>(F + 1), 0.0);
  for (int a = 0; a < g.num_assets; ++a) {
    const int r = m.assetRate[static_cast<std::size_t>(a)];
    if (r > 0 && m.fxOf[static_cast<std::size_t>(r)] < 0) m.fxOf[static_cast<std::size_t>(r)] = a;
  }
  for (int r = 1; r <= F; ++r) {
    const int a = m.fxOf[static_cast<std::size_t>(r)];
    if (a >= 0)
      m.quanto[static_cast<std::size_t>(r)] =
          -corr[static_cast<std::size_t>(r * n + (1 + F + a))] * m.rates[static_cast<std::size_t>(r)].s *
          m.vol[static_cast<std::size_t>(a)];
  }
  return m;
}

Trade makeTrade(const ccr_trade& d, const Models& m) {
  Trade t;
  t.kind = d.kind;
  if (d.kind == CCR_TRADE_SWAP) {
    if (!d.schedule || d.schedule_size < 2) invalid("InterestRateSwap: schedule needs at least two dates");
    t.sched.assign(d.schedule, d.schedule + d.schedule_size);
    checkSchedule(t.sched);
    if (d.float_schedule && d.float_schedule_size > 0) {
      t.fsched.assign(d.float_schedule, d.float_schedule + d.float_schedule_size);
      checkSchedule(t.fsched);
    }
    t.sign = d.pay_fixed ? 1.0 : -1.0;
    t.notional = d.notional;
    t.K = d.fixed_rate;
    t.maturity = std::max(t.sched.back(), t.fsched.empty() ? t.sched.back() : t.fsched.back());
    if (d.currency < 0 || d.currency >= static_cast<int>(m.rates.size()))
      fail(CCR_ERR_OUT_OF_RANGE, "InterestRateSwap: unknown currency");
    t.rate = d.currency;
    if (t.rate > 0 && m.fxOf[static_cast<std::size_t>(t.rate)] < 0)
      invalid("InterestRateSwap: no FX asset converts the swap currency");
  } else if (d.kind == CCR_TRADE_ASSET_FORWARD) {
    if (d.maturity <= 0.0) invalid("AssetForward: maturity must be positive");
    if (d.asset < 0 || d.asset >= static_cast<int>(m.spot.size()))
      fail(CCR_ERR_OUT_OF_RANGE, "ScenarioSet: unknown asset");
    t.asset = d.asset;
    t.notional = d.notional;
    t.K = d.fixed_rate;
    t.maturity = d.maturity;
    const int r = m.assetRate[static_cast<std::size_t>(d.asset)];
    if (r > 0 && m.fxOf[static_cast<std::size_t>(r)] != d.asset)
      fail(CCR_ERR_UNSUPPORTED, "AssetForward: two assets convert the same foreign rate");
  } else if (d.kind == CCR_TRADE_BERMUDAN_SWAPTION) {
    // BermudanSwaption ctor: src/instruments/bermudan_swaption.cpp:17-32.
    if (!d.schedule || d.schedule_size < 2) invalid("InterestRateSwap: schedule needs at least two dates");
    t.sched.assign(d.schedule, d.schedule + d.schedule_size);
    checkSchedule(t.sched);
    if (d.float_schedule && d.float_schedule_size > 0) invalid("BermudanSwaption: the underlying has a single schedule");
    if (d.currency != 0) fail(CCR_ERR_UNSUPPORTED, "BermudanSwaption: domestic underlying only");
    if (!d.exercise_times || d.num_exercises <= 0) invalid("BermudanSwaption: no exercise dates");
    if (d.num_exercises > kMaxExercises) fail(CCR_ERR_UNSUPPORTED, "BermudanSwaption: at most 64 exercise dates");
    t.exercise.assign(d.exercise_times, d.exercise_times + d.num_exercises);
    std::sort(t.exercise.begin(), t.exercise.end());
    for (double e : t.exercise) {
      const bool on = std::any_of(t.sched.begin(), t.sched.end() - 1, [e](double x) { return std::fabs(x - e) < 1e-8; });
      if (!on) invalid("BermudanSwaption: exercise dates must be start dates of the underlying periods");
    }
    if (d.regression_degree < 0 || d.regression_degree > 3)
      fail(CCR_ERR_UNSUPPORTED, "BermudanSwaption: regression degree 0..3");
    t.sign = d.pay_fixed ? 1.0 : -1.0;
    t.notional = d.notional;
    t.K = d.fixed_rate;
    t.maturity = t.sched.back();
    t.rate = 0;
    t.cash = d.cash_settled != 0;
    t.position = d.position;
    t.degree = d.regression_degree;
  } else if (d.kind == CCR_TRADE_EUROPEAN_OPTION) {
    // EuropeanOption ctor: src/instruments/european_option.cpp:9-18.
    if (d.maturity <= 0.0 || d.fixed_rate <= 0.0) invalid("EuropeanOption: invalid expiry or strike");
    if (d.asset < 0 || d.asset >= static_cast<int>(m.spot.size()))
      fail(CCR_ERR_OUT_OF_RANGE, "ScenarioSet: unknown asset");
    if (m.assetRate[static_cast<std::size_t>(d.asset)] != 0)
      fail(CCR_ERR_UNSUPPORTED, "EuropeanOption: the asset needs a deterministic carry curve");
    if (d.option_type != CCR_OPTION_CALL && d.option_type != CCR_OPTION_PUT) invalid("EuropeanOption: unknown type");
    t.asset = d.asset;
    t.notional = d.notional;
    t.K = d.fixed_rate;
    t.maturity = d.maturity;
    t.optionType = d.option_type;
  } else {
    invalid("unknown trade kind");
  }
  return t;
}

// Terms of one date before A/B are attached.
struct Collector {
  std::map<std::pair<int, double>, double> bond;  // (rate, T) -> coefficient
  std::map<std::tuple<int, int, double, double>, double> fix;  // (rate, k, ts, te) -> coefficient
  double c0[kMaxRates] = {0, 0, 0, 0};
  double lin[kMaxAssets] = {0, 0, 0, 0};
  std::vector<const Trade*> options;  // European options alive at t
  double t = 0.0;
  void addBond(int r, double T, double c) {
    if (T <= t) {
      c0[r] += c;  // HullWhite1F::zeroBond returns 1 for T <= t
    } else {
      bond[{r, T}] += c;
    }
  }
};

void collectSwapPeriods(const Trade& tr, const TimeGrid& grid, Collector& col) {
  const double t = col.t, sN = tr.sign * tr.notional;
  const int r = tr.rate;
  auto floating = [&](double ts, double te, bool withFixed) {
    const double tau = te - ts;
    const double fixedPart = withFixed ? tr.K * tau : 0.0;
    if (ts >= t - kPaidTolerance) {
      col.addBond(r, ts, sN);
      col.addBond(r, te, -sN * (1.0 + fixedPart));
    } else {
      const int k = static_cast<int>(grid.indexAtOrBefore(ts));
      col.fix[{r, k, ts, te}] += sN;
      col.addBond(r, te, -sN * (1.0 + fixedPart));
    }
  };
  if (tr.fsched.empty()) {
    for (std::size_t i = 1; i < tr.sched.size(); ++i) {
      const double ts = tr.sched[i - 1], te = tr.sched[i];
      if (te <= t + kPaidTolerance) continue;
      floating(ts, te, true);
    }
    return;
  }
  for (std::size_t i = 1; i < tr.fsched.size(); ++i) {
    const double ts = tr.fsched[i - 1], te = tr.fsched[i];
    if (te <= t + kPaidTolerance) continue;
    floating(ts, te, false);
  }
  for (std::size_t i = 1; i < tr.sched.size(); ++i) {
    const double ts = tr.sched[i - 1], te = tr.sched[i];
    if (te <= t + kPaidTolerance) continue;
    col.addBond(r, te, -sN * tr.K * (te - ts));
  }
}

void collect(const Trade& tr, const Models& m, const TimeGrid& grid, Collector& col) {
  const double t = col.t;
  if (tr.kind == CCR_TRADE_SWAP) {
    if (t >= tr.maturity - kPaidTolerance) return;
    collectSwapPeriods(tr, grid, col);
    return;
  }
  if (t >= tr.maturity - kPaidTolerance) return;
  if (tr.kind == CCR_TRADE_EUROPEAN_OPTION) {
    col.options.push_back(&tr);
    return;
  }
  // AssetForward: N S Q(t,T) - N K P(t,T).
  col.addBond(0, tr.maturity, -tr.notional * tr.K);
  const int r = m.assetRate[static_cast<std::size_t>(tr.asset)];
  if (r == 0) {
    const Curve& c = m.carry[static_cast<std::size_t>(tr.asset)];
    col.lin[tr.asset] += tr.notional * (c.discount(tr.maturity) / c.discount(t));
  } else {
    col.addBond(r, tr.maturity, tr.notional);
  }
}

// The valued dates of a netting set on `sim` (reporting dates of `rep`, margin-call dates of a
// CSA, or every date) and, per (rate, fixing date), the last valued date whose fixed floating
// coupons read that fixing (the condition of collectSwapPeriods' fixed branch).
void fixingUses(const std::vector<Trade>& trades, const ccr_netting_set& ns, const TimeGrid& rep, const TimeGrid& sim,
                bool value_all, std::map<std::pair<int, int>, int>& lastUse) {
  const int T = static_cast<int>(sim.size());
  std::vector<char> valued(static_cast<std::size_t>(T), value_all ? 1 : 0);
  for (double t : rep.t)
    if (const auto j = sim.find(t)) valued[*j] = 1;
  if (ns.has_collateral)
    for (int j = 0; j < T; ++j)
      valued[sim.indexAtOrBefore(std::max(sim[static_cast<std::size_t>(j)] - ns.collateral.margin_period_of_risk, 0.0))] = 1;
  for (int j = 0; j < T; ++j) {
    if (!valued[static_cast<std::size_t>(j)]) continue;
    const double t = sim[static_cast<std::size_t>(j)];
    for (const Trade& tr : trades) {
      if (tr.kind != CCR_TRADE_SWAP || t >= tr.maturity - kPaidTolerance) continue;
      const std::vector<double>& fl = tr.fsched.empty() ? tr.sched : tr.fsched;
      for (std::size_t i = 1; i < fl.size(); ++i) {
        const double ts = fl[i - 1], te = fl[i];
        if (te <= t + kPaidTolerance || ts >= t - kPaidTolerance) continue;
        const auto key = std::make_pair(tr.rate, static_cast<int>(sim.indexAtOrBefore(ts)));
        auto it = lastUse.find(key);
        if (it == lastUse.end() || it->second < j) lastUse[key] = j;
      }
    }
  }
}

// Fixing slots by interval colouring of [fixing date, last use]: slotOf (rate, k) and, per date
// k, the (rate, slot) pairs stored there; returns the slot count.
int assignSlots(const std::map<std::pair<int, int>, int>& lastUse, int T, std::map<std::pair<int, int>, int>& slotOf,
                std::vector<std::vector<std::pair<int, int>>>& storeAt) {
  std::vector<int> slotFreeAfter;  // last use of the interval currently in each slot
  storeAt.assign(static_cast<std::size_t>(T), {});
  std::vector<std::tuple<int, int, int>> intervals;  // (k, rate, last)
  for (const auto& kv : lastUse) intervals.emplace_back(kv.first.second, kv.first.first, kv.second);
  std::sort(intervals.begin(), intervals.end());
  for (const auto& iv : intervals) {
    const int k = std::get<0>(iv), r = std::get<1>(iv), last = std::get<2>(iv);
    int slot = -1;
    for (std::size_t q = 0; q < slotFreeAfter.size(); ++q)
      if (slotFreeAfter[q] < k) {
        slot = static_cast<int>(q);
        break;
      }
    if (slot < 0) {
      slot = static_cast<int>(slotFreeAfter.size());
      slotFreeAfter.push_back(0);
    }
    slotFreeAfter[static_cast<std::size_t>(slot)] = last;
    slotOf[{r, k}] = slot;
    storeAt[static_cast<std::size_t>(k)].emplace_back(r, slot);
  }
  return static_cast<int>(slotFreeAfter.size());
}

// Bound on |x_r(t)| used by the bond polynomial: |mean| + kPolySigmas standard deviations of the
// Gaussian Hull-White state (constant drift `mu` under the simulation measure).
double stateBound(const HullWhite& hw, double mu, double t) {
  const double a = hw.a, sd = hw.s * std::sqrt((1.0 - std::exp(-2.0 * a * t)) / (2.0 * a));
  return std::fabs(mu) * (1.0 - std::exp(-a * t)) / a + kPolySigmas * sd + 1e-12;
}

}  // namespace

// Power-series form of a date's merged bond terms (cA_i, B_i), i < n, plus the constant c0:
// coef[k] = [k == 0] c0 + sum_i cA_i (-B_i)^k / k!, truncated at the first degree K whose
// remainder bound sum_i |cA_i| (B_i X)^{K+1} / (K+1)! e^{B_i X} is below kPolyTolerance times
// sum_i |cA_i| for |x| <= X. Returns false (no polynomial) when no degree up to kPolyMaxDegree
// reaches the tolerance. Accumulated in long double.
bool bondPolynomial(const double* terms, int n, double c0, double xmax, std::vector<double>& coef, double tolerance) {
  coef.clear();
  long double gross = 0.0L, bmax = 0.0L;
  for (int i = 0; i < n; ++i) {
    gross += std::fabs(static_cast<long double>(terms[2 * i]));
    bmax = std::max(bmax, static_cast<long double>(std::fabs(terms[2 * i + 1])));
  }
  if (n == 0) {
    coef.push_back(c0);
    return true;
  }
  std::vector<long double> pw(static_cast<std::size_t>(n));  // cA_i (-B_i)^k / k!
  std::vector<long double> tail(static_cast<std::size_t>(n)); // |cA_i| (B_i X)^{k+1} / (k+1)! e^{B_i X}
  for (int i = 0; i < n; ++i) {
    pw[static_cast<std::size_t>(i)] = terms[2 * i];
    const long double bx = std::fabs(static_cast<long double>(terms[2 * i + 1])) * xmax;
    tail[static_cast<std::size_t>(i)] = std::fabs(static_cast<long double>(terms[2 * i])) * bx * std::exp(bx);
  }
  for (int k = 0; k <= kPolyMaxDegree; ++k) {
    long double ck = k == 0 ? static_cast<long double>(c0) : 0.0L, rest = 0.0L;
    for (int i = 0; i < n; ++i) {
      ck += pw[static_cast<std::size_t>(i)];
      rest += tail[static_cast<std::size_t>(i)];
      pw[static_cast<std::size_t>(i)] *= -static_cast<long double>(terms[2 * i + 1]) / (k + 1);
      tail[static_cast<std::size_t>(i)] *= std::fabs(static_cast<long double>(terms[2 * i + 1])) * xmax / (k + 2);
    }
    coef.push_back(static_cast<double>(ck));
    if (rest <= static_cast<long double>(tolerance) * gross) return true;
  }
  coef.clear();
  return false;
}

TimeGrid reportingGrid(const TimeGrid& base, const ccr_netting_set& set) {
  std::vector<double> ev;
  for (int i = 0; i < set.num_trades; ++i) {
    const auto e = eventTimes(set.trades[i]);
    ev.insert(ev.end(), e.begin(), e.end());
  }
  return base.merged(ev);
}

TimeGrid simulationGrid(const TimeGrid& base, const ccr_netting_set& set) {
  TimeGrid g = reportingGrid(base, set);
  if (set.has_collateral) g = g.lagged(set.collateral.margin_period_of_risk);
  return g;
}

void generateStream(uint64_t seed, int64_t n, double* out) {
  std::mt19937_64 engine(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i) out[i] = normal(engine);
}

TimeGrid bookReportingGrid(const TimeGrid& base, const ccr_netting_set* sets, int32_t num_sets) {
  std::vector<double> r;
  for (int s = 0; s < num_sets; ++s) {
    const TimeGrid g = reportingGrid(base, sets[s]);
    r.insert(r.end(), g.t.begin(), g.t.end());
  }
  return TimeGrid(r);
}

TimeGrid bookSimulationGrid(const TimeGrid& rep, const ccr_netting_set* sets, int32_t num_sets) {
  std::vector<double> v;
  for (int s = 0; s < num_sets; ++s) {
    const TimeGrid g = simulationGrid(rep, sets[s]);
    v.insert(v.end(), g.t.begin(), g.t.end());
  }
  return TimeGrid(v);
}

namespace {

// Continuation GEMM tiles of a group of swaptions (global indices, all valued on the dates vj
// with continuation dates vcd): runs of dates on which every swaption of the group has the same
// next exercise (an epoch), kGemmDates dates (kNarrowDates in narrow tiles, for groups of at most
// kNarrowCols) x up to kGemmCols of the swaptions that still have an exercise ahead.
void buildTiles(HostPlan& plan, const std::vector<int>& group, const std::vector<int>& vj,
                const std::vector<int>& vcd) {
  if (group.empty() || vj.empty()) return;
  auto next = [&](int b, int j) { return plan.bdates[static_cast<std::size_t>(plan.berms[static_cast<std::size_t>(b)].date_rec + j)].next_k; };
  auto sameEpoch = [&](int ja, int jb) {
    for (int b : group)
      if (next(b, ja) != next(b, jb)) return false;
    return true;
  };
  const bool narrow = static_cast<int>(group.size()) <= kNarrowCols;
  const std::size_t maxDates = static_cast<std::size_t>(narrow ? kNarrowDates : kGemmDates);
  std::size_t i0 = 0;
  while (i0 < vj.size()) {
    std::size_t i1 = i0 + 1;
    while (i1 < vj.size() && i1 - i0 < maxDates && sameEpoch(vj[i0], vj[i1])) ++i1;
    std::vector<int> live;
    for (int b : group)
      if (next(b, vj[i0]) < plan.berms[static_cast<std::size_t>(b)].ne) live.push_back(b);
    for (std::size_t c0 = 0; c0 < live.size(); c0 += kGemmCols) {
      GemmTile tile{};
      tile.d_off = static_cast<int32_t>(plan.tcd.size());
      tile.nd = static_cast<int32_t>(i1 - i0);
      for (std::size_t i = i0; i < i1; ++i) plan.tcd.push_back(vcd[i]);
      tile.nb = static_cast<int32_t>(std::min<std::size_t>(kGemmCols, live.size() - c0));
      tile.col_off = static_cast<int32_t>(plan.gcols.size());
      for (std::size_t c = c0; c < c0 + static_cast<std::size_t>(tile.nb); ++c) {
        const BermDev& bm = plan.berms[static_cast<std::size_t>(live[c])];
        plan.gcols.push_back(bm.ev_off + next(live[c], vj[i0]));
        plan.gberm.push_back(live[c]);
      }
      (narrow ? plan.ntiles : plan.tiles).push_back(tile);
    }
    i0 = i1;
  }
}

}  // namespace

HostPlan buildPlan(const ccr_scenario_generator& gen, const ccr_simulation_config& cfg, const double* base_grid,
                   int32_t base_size, const ccr_netting_set* sets, int32_t num_sets, double pfe_quantile,
                   int32_t precision, int32_t block, bool per_trade, bool book, bool value_all,
                   const std::vector<double>* book_rep, const std::vector<double>* book_sim, int32_t basis,
                   int32_t nested_inner, double nested_mpor) {
  if (!sets || num_sets <= 0) invalid("no netting sets");
  if (nested_inner > 0 && (book || per_trade))
    fail(CCR_ERR_UNSUPPORTED, "nested initial margin: not with book or per-trade plans");
  if (nested_inner > 0 && !(nested_mpor > 0.0)) invalid("nested initial margin: the margin period must be positive");
  if (cfg.num_paths <= 0) invalid("ScenarioGenerator::generate: numPaths must be positive");
  if (cfg.num_paths > (int64_t(1) << 31) - 1) fail(CCR_ERR_UNSUPPORTED, "at most 2^31 - 1 paths");
  if (pfe_quantile < 0.0 || pfe_quantile > 1.0) invalid("quantile: level must be in [0, 1]");
  if (block <= 0 || block % 64 != 0 || block > kMaxBlock) invalid("block size must be a multiple of 64 up to 256");
  const Models m = makeModels(gen);
  const double polyTol = precision == CCR_FP32 ? kPolyToleranceFp32 : kPolyTolerance;
  const TimeGrid base(std::vector<double>(base_grid, base_grid + std::max(0, base_size)));
  HostPlan plan;
  plan.nested_inner = nested_inner > 0 ? nested_inner : 0;
  plan.basis = basis;
  plan.precision = precision;
  plan.pfe_q = pfe_quantile;
  plan.block = block;
  GlobalDev& g = plan.g;
  const int P = static_cast<int>(cfg.num_paths);
  g.num_paths = P;
  g.antithetic = cfg.antithetic ? 1 : 0;
  g.nz = m.nz;
  g.ncorr = m.ncorr;
  g.nrates = static_cast<int>(m.rates.size());
  g.nassets = static_cast<int>(m.spot.size());
  g.ncredits = static_cast<int>(m.credits.size());
  g.nsets = num_sets;
  for (int r = 0; r < kMaxRates; ++r) g.fx_of[r] = r < g.nrates ? m.fxOf[static_cast<std::size_t>(r)] : -1;
  for (int a = 0; a < kMaxAssets; ++a) {
    g.asset_rate[a] = a < g.nassets ? m.assetRate[static_cast<std::size_t>(a)] : 0;
    g.spot[a] = a < g.nassets ? m.spot[static_cast<std::size_t>(a)] : 0.0;
  }
  for (int c = 0; c < kMaxCredits; ++c) {
    const bool on = c < g.ncredits;
    g.credit_off[c] = on ? m.creditOff[static_cast<std::size_t>(c)] : 0;
    g.credit_sub[c] = on ? m.credits[static_cast<std::size_t>(c)].sub : 1;
    g.cir_kappa[c] = on ? m.credits[static_cast<std::size_t>(c)].kappa : 0.0;
    g.cir_theta[c] = on ? m.credits[static_cast<std::size_t>(c)].theta : 0.0;
    g.cir_xi[c] = on ? m.credits[static_cast<std::size_t>(c)].xi : 0.0;
    g.cir_y0[c] = on ? m.credits[static_cast<std::size_t>(c)].y0 : 0.0;
  }
  g.foreign_off = m.foreignOff;
  // Cholesky rows are stored at fixed factor slots (rates 0.., assets kMaxRates.., credits
  // kMaxRates + kMaxAssets..) so the kernel addresses every driver with a compile-time index;
  // columns keep the order of the drawn normals.
  const int nRates = static_cast<int>(m.rates.size()), nAssets = static_cast<int>(m.spot.size());
  for (int i = 0; i < m.ncorr; ++i) {
    const int slot = i < nRates ? i : i < nRates + nAssets ? kMaxRates + (i - nRates) : kMaxRates + kMaxAssets + (i - nRates - nAssets);
    for (int k = 0; k < m.ncorr; ++k)
      g.chol[slot * kMaxCorr + k] = m.chol[static_cast<std::size_t>(i * m.ncorr + k)];
  }

  // Book: every netting set reports on the union of the sets' reporting grids and simulates the
  // union of their simulation grids (each built on that reporting union), so the scenarios are
  // pathwise identical across the sets and additive statistics add up exactly
  // (src/app/analyses.cpp:1690-1709, 1727-1766).
  TimeGrid bookRep, bookSim;
  if (book) {
    bookRep = book_rep ? TimeGrid(*book_rep) : bookReportingGrid(base, sets, num_sets);
    bookSim = book_sim ? TimeGrid(*book_sim) : bookSimulationGrid(bookRep, sets, num_sets);
    // Grids of a larger book: every date this plan's sets need must be on them.
    if (book_rep || book_sim) {
      const TimeGrid ownRep = bookReportingGrid(base, sets, num_sets);
      for (double t : ownRep.t)
        if (!bookRep.find(t)) invalid("book grids: a reporting date of the netting sets is missing");
      for (double t : bookSimulationGrid(bookRep, sets, num_sets).t)
        if (!bookSim.find(t)) invalid("book grids: a simulation date of the netting sets is missing");
    }
  }
  // Book plans: one states pass and one moment / Gram per date serve every set with Bermudans.
  int64_t bookStateOff = -1, bookDomOff = -1;
  std::map<int, int32_t> bookCd;  // simulation date -> continuation date holding its moments
  std::map<std::pair<int, int>, int> bookSlotOf;
  std::vector<std::vector<std::pair<int, int>>> bookStoreAt;
  int bookSlots = 0;
  if (book) {
    std::map<std::pair<int, int>, int> uses;
    for (int s = 0; s < num_sets; ++s) {
      std::vector<Trade> trades;
      for (int i = 0; i < sets[s].num_trades && sets[s].trades; ++i) {
        Trade t = makeTrade(sets[s].trades[i], m);
        if (t.kind != CCR_TRADE_BERMUDAN_SWAPTION) trades.push_back(std::move(t));
      }
      fixingUses(trades, sets[s], bookRep, bookSim, value_all, uses);
    }
    bookSlots = assignSlots(uses, static_cast<int>(bookSim.size()), bookSlotOf, bookStoreAt);
  }
  int64_t maxStride = 0;
  std::vector<RegRec> exRecs, contRecs;  // regression-engine records (basis >= 0)
  // Book plans: swaptions of sets valued on the same dates (their valued-date flags), and those dates.
  std::map<std::vector<char>, std::pair<std::vector<int>, std::vector<int>>> bookGroups;
  for (int s = 0; s < num_sets; ++s) {
    const ccr_netting_set& ns = sets[s];
    if (ns.num_trades <= 0 || !ns.trades) invalid("ExposureEngine::run: empty netting set");
    if (ns.counterparty_credit >= g.ncredits || ns.own_credit >= g.ncredits)
      fail(CCR_ERR_OUT_OF_RANGE, "credit index out of range");
    std::vector<Trade> trades, berms;
    for (int i = 0; i < ns.num_trades; ++i) {
      Trade t = makeTrade(ns.trades[i], m);
      (t.kind == CCR_TRADE_BERMUDAN_SWAPTION ? berms : trades).push_back(std::move(t));
    }
    const TimeGrid rep = book ? bookRep : reportingGrid(base, ns);
    TimeGrid sim = book ? bookSim : simulationGrid(base, ns);
    if (nested_inner > 0) {
      // The inner paths run from every reporting date t to t + mpor, so t + mpor joins the
      // simulation grid up to its horizon (src/gpu/fused_exposure.cpp:199-212).
      if (!berms.empty()) fail(CCR_ERR_UNSUPPORTED, "nested initial margin: Bermudan swaptions are not supported");
      std::vector<double> ahead;
      for (double t : rep.t)
        if (t + nested_mpor <= sim.t.back() + 1e-10) ahead.push_back(t + nested_mpor);
      sim = sim.merged(ahead);
    }
    const int T = static_cast<int>(sim.size()), R = static_cast<int>(rep.size());
    plan.rep_times.push_back(rep.t);
    plan.sim_times.push_back(sim.t);

    SetDev sd{};
    sd.num_dates = T;
    sd.num_rep = R;
    sd.date_off = static_cast<int32_t>(plan.dates.size());
    sd.bond_off = static_cast<int32_t>(plan.bonds.size() / 2);
    sd.fixed_off = static_cast<int32_t>(plan.fixed.size() / 4);
    sd.fixstore_off = static_cast<int32_t>(plan.fixstore.size() / 2);
    sd.poly_off = static_cast<int32_t>(plan.poly.size());
    sd.opt_off = static_cast<int32_t>(plan.opts.size() / kOptStride);
    sd.stream_stride = static_cast<int64_t>(T - 1) * m.nz;
    maxStride = std::max(maxStride, sd.stream_stride);
    sd.cpty = ns.counterparty_credit;
    sd.own = ns.own_credit;

    // Reporting slots and margin-call look-back (collateral.cpp:23-24).
    std::vector<int> repOf(static_cast<std::size_t>(T), -1);
    for (int r = 0; r < R; ++r) {
      const auto j = sim.find(rep[static_cast<std::size_t>(r)]);
      if (!j) fail(CCR_ERR_RUNTIME, "reporting date missing from the simulation grid");
      repOf[*j] = r;
    }
    std::vector<int> call(static_cast<std::size_t>(T), 0);
    std::vector<char> isCall(static_cast<std::size_t>(T), 0);
    std::vector<int> ringPos(static_cast<std::size_t>(T), 0);  // margin ring slot of each call date
    int ring = 1;
    if (ns.has_collateral) {
      const ccr_collateral& c = ns.collateral;
      if (c.threshold_counterparty < 0.0 || c.threshold_own < 0.0 || c.minimum_transfer_amount < 0.0 ||
          c.margin_period_of_risk < 0.0 || c.haircut < 0.0 || c.haircut >= 1.0)
        invalid("simulateCollateral: invalid collateral agreement");
      if (c.type == CCR_COLLATERAL_BOND && c.bond_maturity < sim.t.back() - 1e-10)
        invalid("the collateral bond must mature after the horizon");
      if (c.type == CCR_COLLATERAL_ASSET && (c.asset_index < 0 || c.asset_index >= g.nassets))
        fail(CCR_ERR_OUT_OF_RANGE, "collateral asset index out of range");
      for (int j = 0; j < T; ++j) {
        call[static_cast<std::size_t>(j)] =
            static_cast<int>(sim.indexAtOrBefore(std::max(sim[static_cast<std::size_t>(j)] - c.margin_period_of_risk, 0.0)));
        isCall[static_cast<std::size_t>(call[static_cast<std::size_t>(j)])] = 1;
      }
      // Ring slots by interval colouring: call date c holds its value from c to its last reader
      // (the last j with call[j] = c); a slot is reused by a later call date once its interval
      // has ended (strictly before: a date stores its own value before it reads its call value).
      std::vector<int> lastRead(static_cast<std::size_t>(T), -1);
      for (int j = 0; j < T; ++j) lastRead[static_cast<std::size_t>(call[static_cast<std::size_t>(j)])] = j;
      std::vector<int> freeAfter;  // last reader of the interval currently in each slot
      for (int cd = 0; cd < T; ++cd) {
        if (!isCall[static_cast<std::size_t>(cd)]) continue;
        int slot = -1;
        for (std::size_t q = 0; q < freeAfter.size(); ++q)
          if (freeAfter[q] < cd) {
            slot = static_cast<int>(q);
            break;
          }
        if (slot < 0) {
          slot = static_cast<int>(freeAfter.size());
          freeAfter.push_back(0);
        }
        freeAfter[static_cast<std::size_t>(slot)] = lastRead[static_cast<std::size_t>(cd)];
        ringPos[static_cast<std::size_t>(cd)] = slot;
      }
      ring = std::max(1, static_cast<int>(freeAfter.size()));
      sd.has_csa = 1;
      sd.sec_type = c.type;
      sd.csa_asset = c.asset_index;
      sd.thr_cpty = c.threshold_counterparty;
      sd.thr_own = std::isfinite(c.threshold_own) ? c.threshold_own : -1.0;
      sd.mta = c.minimum_transfer_amount;
      sd.ia = c.independent_amount;
      sd.keep = c.type == CCR_COLLATERAL_CASH ? 1.0 : 1.0 - c.haircut;
    }
    sd.ring_len = ring;

    // Valuation terms per valued date.
    std::vector<Collector> cols(static_cast<std::size_t>(T));
    std::vector<char> valued(static_cast<std::size_t>(T), 0);
    std::map<std::pair<int, int>, int> lastUse;  // (rate, fixing date) -> last date using it
    // Nested initial margin: the date t + mpor of each reporting date (found within 1e-8, as
    // fused_exposure.cpp:612-620), valued for the inner paths.
    std::vector<int32_t> target(static_cast<std::size_t>(T), -1);
    if (nested_inner > 0)
      for (int j = 0; j < T; ++j) {
        if (repOf[static_cast<std::size_t>(j)] < 0) continue;
        if (const auto k = sim.find(sim[static_cast<std::size_t>(j)] + nested_mpor, 1e-8)) {
          if (static_cast<int>(*k) <= j) invalid("nested initial margin: the margin period must be positive");
          target[static_cast<std::size_t>(j)] = static_cast<int32_t>(*k);
        }
      }
    std::vector<char> isTarget(static_cast<std::size_t>(T), 0);
    for (int32_t k : target)
      if (k >= 0) isTarget[static_cast<std::size_t>(k)] = 1;
    for (int j = 0; j < T; ++j) {
      valued[static_cast<std::size_t>(j)] = value_all || repOf[static_cast<std::size_t>(j)] >= 0 ||
                                            isCall[static_cast<std::size_t>(j)] || isTarget[static_cast<std::size_t>(j)];
      if (!valued[static_cast<std::size_t>(j)]) continue;
      Collector& col = cols[static_cast<std::size_t>(j)];
      col.t = sim[static_cast<std::size_t>(j)];
      for (const Trade& tr : trades) collect(tr, m, sim, col);
      for (const auto& kv : col.fix) {
        const auto key = std::make_pair(std::get<0>(kv.first), std::get<1>(kv.first));
        auto it = lastUse.find(key);
        if (it == lastUse.end() || it->second < j) lastUse[key] = j;
      }
    }
    // Bermudan swaptions: exercise dates on the grid, the next exercise per date, the underlying's
    // value terms where an exercised path holds it (bermudan_swaption.cpp:55-62, 104-140).
    const int nB = static_cast<int>(berms.size());
    std::vector<std::vector<int>> bEidx(static_cast<std::size_t>(nB)), bNext(static_cast<std::size_t>(nB));
    std::vector<std::vector<Collector>> bCols(static_cast<std::size_t>(nB));
    std::vector<std::vector<char>> bHas(static_cast<std::size_t>(nB));
    for (int b = 0; b < nB; ++b) {
      const Trade& tr = berms[static_cast<std::size_t>(b)];
      for (double e : tr.exercise) {
        const auto idx = sim.find(e, 1e-8);
        if (!idx) invalid("BermudanSwaption: exercise date not on the simulation grid");
        bEidx[static_cast<std::size_t>(b)].push_back(static_cast<int>(*idx));
      }
      Trade und = tr;
      und.kind = CCR_TRADE_SWAP;
      auto& nxt = bNext[static_cast<std::size_t>(b)];
      auto& bc = bCols[static_cast<std::size_t>(b)];
      auto& has = bHas[static_cast<std::size_t>(b)];
      nxt.assign(static_cast<std::size_t>(T), 0);
      bc.resize(static_cast<std::size_t>(T));
      has.assign(static_cast<std::size_t>(T), 0);
      for (int j = 0; j < T; ++j) {
        const double t = sim[static_cast<std::size_t>(j)];
        int k = 0;
        while (k < static_cast<int>(tr.exercise.size()) && tr.exercise[static_cast<std::size_t>(k)] <= t + kPaidTolerance) ++k;
        nxt[static_cast<std::size_t>(j)] = k;
        if (!valued[static_cast<std::size_t>(j)] || tr.cash || k == 0) continue;
        Collector& c = bc[static_cast<std::size_t>(j)];
        c.t = t;
        collect(und, m, sim, c);
        has[static_cast<std::size_t>(j)] = 1;
      }
    }
    // Fixing slots: the set's own, or in a book the book's (the fixing states are those of the
    // shared scenarios, one LDS column per path serving every set of a workgroup).
    std::map<std::pair<int, int>, int> slotOf;
    std::vector<std::vector<std::pair<int, int>>> storeAt;
    if (book) {
      slotOf = bookSlotOf;
      storeAt = bookStoreAt;
      sd.num_slots = bookSlots;
    } else {
      sd.num_slots = assignSlots(lastUse, T, slotOf, storeAt);
    }

    // Per-trade plans: every trade of the set gets its own valuation records (its terms alone);
    // the set's records then carry none, and the kernel nets the trades' values in trade order.
    const int G = per_trade ? ns.num_trades : 0;
    if (per_trade) {
      if (ns.num_trades > kMaxTradeGroups)
        fail(CCR_ERR_UNSUPPORTED, "per-trade statistics: at most 16 trades per netting set");
      if (!berms.empty()) fail(CCR_ERR_UNSUPPORTED, "per-trade statistics: Bermudan swaptions are not supported");
      sd.num_groups = G;
      sd.tdate_off = static_cast<int32_t>(plan.tdates.size());
      plan.max_groups = std::max(plan.max_groups, G);
    }

    // The value terms of one collector on date t into rec's term ranges and the date reals
    // (constants, deterministic-carry forwards, state bounds).
    auto emitTerms = [&](const Collector& col, double t, bool isValued, DateRec& rec, double* reals) {
      for (int r = 0; r < kMaxRates; ++r) {
        rec.bond_beg[r] = static_cast<int32_t>(plan.bonds.size() / 2) - sd.bond_off;
        if (r < g.nrates) {
          const HullWhite& hw = m.rates[static_cast<std::size_t>(r)];
          for (auto it = col.bond.lower_bound({r, -1e300}); it != col.bond.end() && it->first.first == r; ++it) {
            const double T_ = it->first.second;
            plan.bonds.push_back(it->second * hw.A(t, T_));
            plan.bonds.push_back(hw.B(t, T_));
          }
        }
      }
      rec.bond_beg[kMaxRates] = static_cast<int32_t>(plan.bonds.size() / 2) - sd.bond_off;
      // Bond polynomials of every rate, interleaved row by row (highest power first, shorter
      // series zero-padded on top); none unless every rate has one.
      rec.poly_beg = static_cast<int32_t>(plan.poly.size()) - sd.poly_off;
      rec.poly_rows = 0;
      if (isValued) {
        std::vector<std::vector<double>> coef(static_cast<std::size_t>(g.nrates));
        bool all = true;
        std::size_t rows = 0;
        for (int r = 0; r < g.nrates; ++r) {
          const int b0 = sd.bond_off + rec.bond_beg[r], b1 = sd.bond_off + rec.bond_beg[r + 1];
          const double xmax =
              stateBound(m.rates[static_cast<std::size_t>(r)], m.quanto[static_cast<std::size_t>(r)], t);
          reals[kDateXmax + r] = xmax;
          auto& c = coef[static_cast<std::size_t>(r)];
          all = all && bondPolynomial(plan.bonds.data() + 2 * b0, b1 - b0, col.c0[r], xmax, c, polyTol);
          rows = std::max(rows, c.size());
        }
        // A record with no bond term and no constant on any rate (no cash flow left on this date)
        // keeps no rows: the kernel's sum of the (empty) terms is the same exact zero.
        bool anyTerm = false;
        for (int r = 0; r < g.nrates; ++r)
          anyTerm = anyTerm || rec.bond_beg[r + 1] > rec.bond_beg[r] || col.c0[r] != 0.0;
        if (all && anyTerm) {
          rows = (rows + kPolyBatch - 1) / kPolyBatch * kPolyBatch;
          for (std::size_t k = 0; k < rows; ++k) {
            const std::size_t power = rows - 1 - k;
            for (const auto& c : coef) plan.poly.push_back(power < c.size() ? c[power] : 0.0);
          }
          rec.poly_rows = static_cast<int32_t>(rows);
          plan.poly_evaluations += static_cast<int64_t>(rows) * g.nrates * P;
        }
      }
      for (int r = 0; r < kMaxRates; ++r) {
        rec.fix_beg[r] = static_cast<int32_t>(plan.fixed.size() / 4) - sd.fixed_off;
        if (r < g.nrates) {
          const HullWhite& hw = m.rates[static_cast<std::size_t>(r)];
          for (const auto& kv : col.fix) {
            if (std::get<0>(kv.first) != r) continue;
            const int k = std::get<1>(kv.first);
            const double ts = std::get<2>(kv.first), te = std::get<3>(kv.first);
            const double tk = sim[static_cast<std::size_t>(k)];
            const double As = ts <= tk ? 1.0 : hw.A(tk, ts), Bs = ts <= tk ? 0.0 : hw.B(tk, ts);
            const double Ae = te <= tk ? 1.0 : hw.A(tk, te), Be = te <= tk ? 0.0 : hw.B(tk, te);
            plan.fixed.push_back(kv.second * (As / Ae) * hw.A(t, te));
            plan.fixed.push_back(hw.B(t, te));
            plan.fixed.push_back(Bs - Be);
            plan.fixed.push_back(static_cast<double>(slotOf.at({r, k})));
          }
        }
      }
      rec.fix_beg[kMaxRates] = static_cast<int32_t>(plan.fixed.size() / 4) - sd.fixed_off;
      // European options (european_option.cpp:20-35): Black on the forward S Q(t,T) / P(t,T).
      rec.opt_beg = static_cast<int32_t>(plan.opts.size() / kOptStride) - sd.opt_off;
      for (const Trade* op : col.options) {
        const HullWhite& hw = m.rates[0];
        const double T_ = op->maturity;
        const Curve& carry = m.carry[static_cast<std::size_t>(op->asset)];
        const double o[kOptStride] = {static_cast<double>(op->asset), op->optionType == CCR_OPTION_CALL ? 1.0 : -1.0,
                                      op->notional, op->K, hw.A(t, T_), hw.B(t, T_),
                                      carry.discount(T_) / carry.discount(t),
                                      m.vol[static_cast<std::size_t>(op->asset)] * std::sqrt(T_ - t)};
        plan.opts.insert(plan.opts.end(), o, o + kOptStride);
      }
      rec.opt_end = static_cast<int32_t>(plan.opts.size() / kOptStride) - sd.opt_off;
      for (int r = 0; r < kMaxRates; ++r) reals[kDateC0 + r] = col.c0[r];
      for (int a = 0; a < kMaxAssets; ++a) reals[kDateLin + a] = col.lin[a];
    };

    // Emit dates.
    int64_t setTerms = 0;
    int32_t valuedSoFar = 0;
    for (int j = 0; j < T; ++j) {
      const double t = sim[static_cast<std::size_t>(j)];
      DateRec rec{};
      rec.rep = repOf[static_cast<std::size_t>(j)];
      rec.flags = (valued[static_cast<std::size_t>(j)] ? kValued : 0) | (rec.rep >= 0 ? kReporting : 0) |
                  (isCall[static_cast<std::size_t>(j)] ? kStoreValue : 0);
      rec.call_pos = ringPos[static_cast<std::size_t>(call[static_cast<std::size_t>(j)])];
      rec.store_pos = isCall[static_cast<std::size_t>(j)] ? ringPos[static_cast<std::size_t>(j)] : -1;
      rec.vidx = valued[static_cast<std::size_t>(j)] ? valuedSoFar++ : -1;
      double reals[kDateRealStride] = {0};
      const int bondBase = static_cast<int>(plan.bonds.size() / 2), fixBase = static_cast<int>(plan.fixed.size() / 4);
      if (per_trade) {
        Collector none;
        none.t = t;
        emitTerms(none, t, false, rec, reals);
        for (int i = 0; i < G; ++i) {
          Collector one;
          one.t = t;
          if (valued[static_cast<std::size_t>(j)]) collect(trades[static_cast<std::size_t>(i)], m, sim, one);
          DateRec trec{};
          trec.flags = rec.flags;
          double treals[kDateRealStride] = {0};
          emitTerms(one, t, valued[static_cast<std::size_t>(j)] != 0, trec, treals);
          plan.tdates.push_back(trec);
          plan.tdreals.insert(plan.tdreals.end(), treals, treals + kDateRealStride);
        }
      } else {
        emitTerms(cols[static_cast<std::size_t>(j)], t, valued[static_cast<std::size_t>(j)] != 0, rec, reals);
      }
      rec.fixstore_beg = static_cast<int32_t>(plan.fixstore.size() / 2) - sd.fixstore_off;
      for (const auto& rs : storeAt[static_cast<std::size_t>(j)]) {
        plan.fixstore.push_back(rs.first);
        plan.fixstore.push_back(rs.second);
      }
      rec.fixstore_end = static_cast<int32_t>(plan.fixstore.size() / 2) - sd.fixstore_off;
      if (ns.has_collateral && ns.collateral.type == CCR_COLLATERAL_BOND) {
        const double T_ = ns.collateral.bond_maturity;
        reals[kDateBond] = T_ <= t ? 1.0 : m.rates[0].A(t, T_);
        reals[kDateBond + 1] = T_ <= t ? 0.0 : m.rates[0].B(t, T_);
      }
      const int nTerms = (static_cast<int>(plan.bonds.size() / 2) - bondBase) + (static_cast<int>(plan.fixed.size() / 4) - fixBase);
      setTerms += nTerms;
      plan.dates.push_back(rec);
      plan.date_reals.insert(plan.date_reals.end(), reals, reals + kDateRealStride);
      plan.nested_target.push_back(target[static_cast<std::size_t>(j)]);

      // Step coefficients t_{j-1} -> t_j.
      double st[kStepStride] = {0};
      if (j > 0) {
        const double t0 = sim[static_cast<std::size_t>(j - 1)], t1 = t, dt = t1 - t0;
        for (int r = 0; r < g.nrates; ++r) {
          const HullWhite& hw = m.rates[static_cast<std::size_t>(r)];
          const HullWhite::Step hs = hw.step(t0, t1);
          const double mu = m.quanto[static_cast<std::size_t>(r)];
          double* o = st + kStepRate + 8 * r;
          o[0] = hs.e;
          o[1] = hs.sdX;
          o[2] = hs.c1;
          o[3] = hs.c2;
          o[4] = hs.xc;
          o[5] = mu != 0.0 ? mu * (1.0 - hs.e) / hw.a : 0.0;
          o[6] = mu != 0.0 ? mu * (dt - (1.0 - hs.e) / hw.a) / hw.a : 0.0;
          o[7] = hs.iphi;
        }
        for (int a = 0; a < g.nassets; ++a) {
          double* o = st + kStepAsset + 3 * a;
          const double vol = m.vol[static_cast<std::size_t>(a)];
          o[0] = m.assetRate[static_cast<std::size_t>(a)] == 0
                     ? std::log(m.carry[static_cast<std::size_t>(a)].discount(t0) / m.carry[static_cast<std::size_t>(a)].discount(t1))
                     : 0.0;
          o[1] = vol * std::sqrt(dt);
          o[2] = 0.5 * vol * vol * dt;
        }
        for (int c = 0; c < g.ncredits; ++c) {
          const CirModel& cm = m.credits[static_cast<std::size_t>(c)];
          double* o = st + kStepCredit + kStepCreditStride * c;
          const double h = dt / static_cast<double>(cm.sub);
          o[0] = std::sqrt(dt);
          o[1] = h;
          o[2] = cm.shift(t1);
          double remT = dt;
          for (int k = 0; k < cm.sub; ++k) {
            if (k + 1 < cm.sub) {
              o[3 + k] = h / remT;
              o[3 + kMaxSub + k] = std::sqrt(h * (remT - h) / remT);
            }
            remT -= h;
          }
        }
        if (dt <= 0.0) invalid("HullWhite1F::evolve: t1 must exceed t0");
      }
      plan.steps.insert(plan.steps.end(), st, st + kStepStride);
      if (valued[static_cast<std::size_t>(j)]) plan.term_evaluations += static_cast<int64_t>(nTerms) * P;
    }
    plan.total_terms += setTerms;

    // Bermudan records: exercise values (tail swap at each exercise date) and, per date, the next
    // exercise and the underlying's value terms.
    auto emitRange = [&](const Collector& col, double t, double xmax, bool allowFixed) {
      TermRange tr{};
      const HullWhite& hw = m.rates[0];
      tr.bond_beg = static_cast<int32_t>(plan.bonds.size() / 2);
      for (const auto& kv : col.bond) {
        plan.bonds.push_back(kv.second * hw.A(t, kv.first.second));
        plan.bonds.push_back(hw.B(t, kv.first.second));
      }
      tr.bond_end = static_cast<int32_t>(plan.bonds.size() / 2);
      tr.fix_beg = static_cast<int32_t>(plan.fixed.size() / 4);
      for (const auto& kv : col.fix) {
        if (!allowFixed) fail(CCR_ERR_UNSUPPORTED, "BermudanSwaption: exercise value with a fixed coupon");
        const int k = std::get<1>(kv.first);
        const double ts = std::get<2>(kv.first), te = std::get<3>(kv.first);
        const double tk = sim[static_cast<std::size_t>(k)];
        const double As = ts <= tk ? 1.0 : hw.A(tk, ts), Bs = ts <= tk ? 0.0 : hw.B(tk, ts);
        const double Ae = te <= tk ? 1.0 : hw.A(tk, te), Be = te <= tk ? 0.0 : hw.B(tk, te);
        plan.fixed.push_back(kv.second * (As / Ae) * hw.A(t, te));
        plan.fixed.push_back(hw.B(t, te));
        plan.fixed.push_back(Bs - Be);
        plan.fixed.push_back(static_cast<double>(k));  // fixing date: a row of the recorded states
      }
      tr.fix_end = static_cast<int32_t>(plan.fixed.size() / 4);
      tr.c0 = col.c0[0];
      tr.xmax = xmax;
      tr.poly_beg = static_cast<int32_t>(plan.poly.size());
      std::vector<double> coef;
      if (bondPolynomial(plan.bonds.data() + 2 * tr.bond_beg, tr.bond_end - tr.bond_beg, tr.c0, xmax, coef, polyTol)) {
        const std::size_t len = (coef.size() + kPolyBatch - 1) / kPolyBatch * kPolyBatch;
        plan.poly.insert(plan.poly.end(), len - coef.size(), 0.0);
        plan.poly.insert(plan.poly.end(), coef.rbegin(), coef.rend());
        tr.poly_len = static_cast<int32_t>(len);
      }
      // Indices relative to the set's tables, like the date records.
      tr.bond_beg -= sd.bond_off;
      tr.bond_end -= sd.bond_off;
      tr.fix_beg -= sd.fixed_off;
      tr.fix_end -= sd.fixed_off;
      tr.poly_beg -= sd.poly_off;
      return tr;
    };
    sd.berm_beg = static_cast<int32_t>(plan.berms.size());
    sd.state_off = -1;
    sd.dom_off = -1;
    if (nB > 0 && book && bookStateOff >= 0) {
      sd.state_off = bookStateOff;  // same grid, same draws: the states of the book's first such set
      sd.dom_off = bookDomOff;
    } else if (nB > 0) {
      sd.state_off = plan.state_len;
      sd.dom_off = plan.dom_len;
      plan.dom_len += static_cast<int64_t>(T - 1) * 2 * (cfg.antithetic ? (static_cast<int64_t>(P) + 1) / 2 : P);
      plan.state_len += static_cast<int64_t>(T) * P;
      if (book) {
        bookStateOff = sd.state_off;
        bookDomOff = sd.dom_off;
        plan.states_set = s;
      }
    }
    // Continuation regressions run on every valued date of a set with Bermudans.
    std::vector<int32_t> cdOf(static_cast<std::size_t>(T), -1);
    sd.cd_base = static_cast<int32_t>(plan.cdates.size() / 2);
    if (nB > 0)
      for (int j = 0; j < T; ++j)
        if (valued[static_cast<std::size_t>(j)]) {
          const int32_t cd = static_cast<int32_t>(plan.cdates.size() / 2);
          cdOf[static_cast<std::size_t>(j)] = cd;
          plan.cdates.push_back(s);
          plan.cdates.push_back(j);
          plan.ctimes.push_back(sim[static_cast<std::size_t>(j)]);
          int32_t canon = cd;
          if (book) {
            const auto it = bookCd.find(j);
            if (it == bookCd.end()) bookCd[j] = cd;
            else canon = it->second;
          }
          plan.cmap.push_back(canon);
          if (canon == cd) plan.cuniq.push_back(cd);
          // The swaptions with a value on this date: an exercise still ahead (continuation), or
          // an underlying with cash flows left that exercised paths hold. The others are 0 on
          // every path and the valuation skips them.
          plan.clive_rng.push_back(static_cast<int32_t>(plan.clive.size()));
          for (int b = 0; b < nB; ++b) {
            const Collector& c = bCols[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)];
            const bool ahead = bNext[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)] <
                               static_cast<int>(berms[static_cast<std::size_t>(b)].exercise.size());
            const bool held = bHas[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)] &&
                              (!c.bond.empty() || !c.fix.empty() || c.c0[0] != 0.0);
            if (ahead || held) {
              // The swaption's date records start after those of the set's earlier swaptions.
              LiveRec lr{};
              lr.b = static_cast<int32_t>(plan.berms.size()) + b;
              lr.rec = static_cast<int32_t>(plan.bdates.size()) + b * T + j;
              lr.ne = static_cast<int32_t>(berms[static_cast<std::size_t>(b)].exercise.size());
              lr.position = berms[static_cast<std::size_t>(b)].position;
              plan.clive.push_back(lr);
            }
          }
          plan.clive_rng.push_back(static_cast<int32_t>(plan.clive.size()));
        }
    for (int b = 0; b < nB; ++b) {
      const Trade& tr = berms[static_cast<std::size_t>(b)];
      const int ne = static_cast<int>(tr.exercise.size());
      BermDev bd{};
      bd.set = s;
      bd.ne = ne;
      bd.degree = tr.degree;
      bd.cash = tr.cash ? 1 : 0;
      bd.position = tr.position;
      bd.ev_rec = static_cast<int32_t>(plan.evterms.size());
      bd.date_rec = static_cast<int32_t>(plan.bdates.size());
      bd.eidx_off = static_cast<int32_t>(plan.eidx.size());
      bd.ev_off = static_cast<int32_t>(plan.ev_len / P);
      for (int k = 0; k < ne; ++k) {
        const int je = bEidx[static_cast<std::size_t>(b)][static_cast<std::size_t>(k)];
        const double te = sim[static_cast<std::size_t>(je)];
        Trade tail = tr;  // InterestRateSwap::tail (interest_rate_swap.cpp:97-103)
        tail.kind = CCR_TRADE_SWAP;
        tail.sched.clear();
        for (double x : tr.sched)
          if (x >= tr.exercise[static_cast<std::size_t>(k)] - kPaidTolerance) tail.sched.push_back(x);
        if (tail.sched.size() < 2) invalid("InterestRateSwap::tail: no period starts after fromTime");
        tail.maturity = tail.sched.back();
        Collector ec;
        ec.t = te;
        collect(tail, m, sim, ec);
        plan.eidx.push_back(je);
        plan.evterms.push_back(emitRange(ec, te, stateBound(m.rates[0], 0.0, te), false));
        plan.exlist.push_back(static_cast<int32_t>(plan.berms.size()));
      }
      for (int j = 0; j < T; ++j) {
        BermDate rec{};
        rec.next_k = bNext[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)];
        rec.has_swap = bHas[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)];
        if (rec.has_swap) {
          const double t = sim[static_cast<std::size_t>(j)];
          rec.swap = emitRange(bCols[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)], t,
                               stateBound(m.rates[0], 0.0, t), true);
        }
        plan.bdates.push_back(rec);
        plan.bdate_berm.push_back(static_cast<int32_t>(plan.berms.size()));
        plan.bdate_cd.push_back(cdOf[static_cast<std::size_t>(j)]);
      }
      // Regression-engine records: exercise k at its date (tail k), and the continuation of every
      // valued date with an exercise ahead (tail of the next exercise). For the swap-rate /
      // annuity basis each carries the annuity sum_i tau_i P(t, T_i) and the float leg
      // P(t, T_start) - P(t, T_n) of its tail as value terms on its date.
      if (basis >= 0) {
        auto tailTerms = [&](int k, int j, RegRec& rr) {
          rr.ann = rr.flt = -1;
          if (basis != kBasisSwapRateAnnuity) return;
          const double t = sim[static_cast<std::size_t>(j)];
          std::vector<double> sched;
          for (double x : tr.sched)
            if (x >= tr.exercise[static_cast<std::size_t>(k)] - kPaidTolerance) sched.push_back(x);
          Collector ann, flt;
          ann.t = flt.t = t;
          for (std::size_t i = 1; i < sched.size(); ++i) ann.addBond(0, sched[i], sched[i] - sched[i - 1]);
          flt.addBond(0, sched.front(), 1.0);
          flt.addBond(0, sched.back(), -1.0);
          const double xmax = stateBound(m.rates[0], 0.0, t);
          rr.ann = static_cast<int32_t>(plan.saterms.size());
          plan.saterms.push_back(emitRange(ann, t, xmax, false));
          rr.flt = static_cast<int32_t>(plan.saterms.size());
          plan.saterms.push_back(emitRange(flt, t, xmax, false));
        };
        const int gb = static_cast<int>(plan.berms.size());
        for (int k = 0; k < ne; ++k) {
          RegRec rr{};
          rr.b = gb;
          rr.j = bEidx[static_cast<std::size_t>(b)][static_cast<std::size_t>(k)];
          rr.k = k;
          rr.kind = 0;
          tailTerms(k, rr.j, rr);
          exRecs.push_back(rr);
        }
        for (int j = 0; j < T; ++j) {
          const int k = bNext[static_cast<std::size_t>(b)][static_cast<std::size_t>(j)];
          if (cdOf[static_cast<std::size_t>(j)] < 0 || k >= ne) continue;
          RegRec rr{};
          rr.b = gb;
          rr.j = j;
          rr.k = k;
          rr.kind = 1;
          rr.pad[0] = bd.date_rec + j;  // its BermDate record
          tailTerms(k, j, rr);
          contRecs.push_back(rr);
        }
      }
      plan.ev_len += static_cast<int64_t>(ne) * P;
      plan.cont_len += static_cast<int64_t>(T) * kContStride;
      plan.max_exercises = std::max(plan.max_exercises, ne);
      plan.berms.push_back(bd);
    }
    sd.berm_end = static_cast<int32_t>(plan.berms.size());
    // Continuation GEMM tiles of the set's swaptions (book plans tile all sets together below).
    if (nB > 0) {
      std::vector<int> group, vj, vcd;
      for (int b = 0; b < nB; ++b) group.push_back(sd.berm_beg + b);
      for (int j = 0; j < T; ++j)
        if (valued[static_cast<std::size_t>(j)]) {
          vj.push_back(j);
          vcd.push_back(cdOf[static_cast<std::size_t>(j)]);
        }
      if (book) {
        bookGroups[valued].first.insert(bookGroups[valued].first.end(), group.begin(), group.end());
        bookGroups[valued].second = vj;
      } else {
        buildTiles(plan, group, vj, vcd);
      }
    }
    plan.revaluations += static_cast<int64_t>(ns.num_trades) * P * R;
    sd.rows_off = static_cast<int64_t>(s) * 0;  // filled below once max_rep is known
    if (nested_inner > 0)
      for (int j = 0; j < T; ++j) {
        const int32_t k = target[static_cast<std::size_t>(j)];
        if (k < 0) continue;
        const int save = plan.dates[static_cast<std::size_t>(sd.date_off + k)].fixstore_end -
                         plan.dates[static_cast<std::size_t>(sd.date_off + j + 1)].fixstore_beg;
        if (save > kNestedMaxSave)
          fail(CCR_ERR_UNSUPPORTED, "nested initial margin: more fixings inside a margin period than supported");
      }
    plan.sets.push_back(sd);
    plan.max_dates = std::max(plan.max_dates, T);
    plan.max_rep = std::max(plan.max_rep, R);
    g.max_ring = std::max(g.max_ring, ring);
    if (ns.has_collateral && ns.collateral.type != CCR_COLLATERAL_CASH) plan.price_rows = std::max(plan.price_rows, ring);
    g.max_slots = std::max(g.max_slots, sd.num_slots);
  }
  // Book tiles: one group per valued-date pattern, its dates' rows from the shared moments.
  for (const auto& kv : bookGroups) {
    std::vector<int> vcd;
    for (int j : kv.second.second) vcd.push_back(bookCd.at(j));
    buildTiles(plan, kv.second.first, kv.second.second, vcd);
  }
  if (basis >= 0 && !plan.berms.empty()) {
    plan.nex_recs = static_cast<int32_t>(exRecs.size());
    plan.rrecs = exRecs;
    plan.rrecs.insert(plan.rrecs.end(), contRecs.begin(), contRecs.end());
    plan.rex.assign(plan.berms.size() * static_cast<std::size_t>(std::max(1, plan.max_exercises)), -1);
    for (int i = 0; i < plan.nex_recs; ++i) {
      const RegRec& rr = plan.rrecs[static_cast<std::size_t>(i)];
      plan.rex[static_cast<std::size_t>(rr.b) * static_cast<std::size_t>(plan.max_exercises) + static_cast<std::size_t>(rr.k)] = i;
    }
    plan.rcont.assign(plan.bdates.size(), -1);
    for (std::size_t i = static_cast<std::size_t>(plan.nex_recs); i < plan.rrecs.size(); ++i)
      plan.rcont[static_cast<std::size_t>(plan.rrecs[i].pad[0])] = static_cast<int32_t>(i);
    if (plan.saterms.empty()) plan.saterms.push_back(TermRange{});
  }
  const int64_t blocksPaths = cfg.antithetic ? (static_cast<int64_t>(P) + 1) / 2 : P;
  plan.stream_len = blocksPaths * maxStride;
  for (int s = 0; s < num_sets; ++s)
    plan.sets[static_cast<std::size_t>(s)].rows_off = static_cast<int64_t>(s) * plan.max_rep * P;
  for (LiveRec& lr : plan.clive) {  // each live record carries its date's BermDate
    const BermDate& bd = plan.bdates[static_cast<std::size_t>(lr.rec)];
    lr.next_k = bd.next_k;
    lr.has_swap = bd.has_swap;
    lr.swap = bd.swap;
  }
  if (plan.clive.empty()) plan.clive.push_back(LiveRec{});  // a valid table even when no swaption is ever live
  // One zero coupon record past the end: the path kernel reads coupons two at a time.
  plan.fixed.insert(plan.fixed.end(), 4, 0.0);
  assignBlocks(plan, block);
  return plan;
}

void assignBlocks(HostPlan& plan, int block) {
  if (block <= 0 || block % 64 != 0 || block > kMaxBlock) invalid("block size must be a multiple of 64 up to 256");
  const int P = plan.g.num_paths, num_sets = static_cast<int>(plan.sets.size());
  const int wavesPerBlock = block / 64;
  const int blocksPerSet = (P + block - 1) / block;
  plan.block = block;
  for (int s = 0; s < num_sets; ++s) {
    SetDev& sd = plan.sets[static_cast<std::size_t>(s)];
    sd.block_begin = s * blocksPerSet;
    sd.num_blocks = blocksPerSet;
    sd.wave_begin = s * blocksPerSet * wavesPerBlock;
    sd.partial_off = static_cast<int64_t>(sd.wave_begin) * plan.max_rep * kStats;
  }
  plan.total_waves = num_sets * blocksPerSet * wavesPerBlock;
  plan.block_set.resize(static_cast<std::size_t>(num_sets * blocksPerSet));
  for (int b = 0; b < num_sets * blocksPerSet; ++b) plan.block_set[static_cast<std::size_t>(b)] = b / blocksPerSet;
}

}  // namespace ccrhip
