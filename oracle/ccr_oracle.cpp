// ccr_oracle.cpp -- FP64 CPU oracle of the exposure path. TEST INFRASTRUCTURE ONLY: imported
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker, never by the
// product library.
//
// A restatement of the reference CPU engine's semantics (paths relative to the reference root):
// every function names the file:line it follows. Floating-point expressions keep the reference's
// operation order so that, on reference-representable inputs, results agree with a build of the
// reference bit for bit or to a few ulps (checked in tests/test_oracle.py against oracle/_ref,
// the reference's own sources compiled by oracle/Makefile).
//
// Extensions beyond the reference model (foreign Hull-White rates, two-leg swaps, pathwise DVA)
// are the oracle's own and are pinned by reductions (sigma_f = 0, equal schedules) and invariants.

#include "ccr_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <optional>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace oracle {

using Vec = std::vector<double>;

// ---------------------------------------------------------------- errors
struct Failure : std::runtime_error {
  int32_t code;
  Failure(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void invalid(const std::string& m) { throw Failure(CCR_ERR_INVALID_ARGUMENT, m); }

// ---------------------------------------------------------------- dense matrix [rows][cols]
struct Mat {
  std::size_t r = 0, c = 0;
  Vec d;
  Mat() = default;
  Mat(std::size_t rows, std::size_t cols, double v = 0.0) : r(rows), c(cols), d(rows * cols, v) {}
  double& operator()(std::size_t i, std::size_t j) { return d[i * c + j]; }
  double operator()(std::size_t i, std::size_t j) const { return d[i * c + j]; }
};

// cholesky: src/core/matrix.cpp:49-68
Mat cholesky(const Mat& a) {
  const std::size_t n = a.r;
  Mat l(n, n, 0.0);
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = 0; j <= i; ++j) {
      if (std::fabs(a(i, j) - a(j, i)) > 1e-10) invalid("cholesky: matrix must be symmetric");
      double s = a(i, j);
      for (std::size_t k = 0; k < j; ++k) s -= l(i, k) * l(j, k);
      if (i == j) {
        if (s <= 0.0) invalid("cholesky: matrix is not positive definite");
        l(i, i) = std::sqrt(s);
      } else {
        l(i, j) = s / l(j, j);
      }
    }
  return l;
}

// Gaussian elimination with partial pivoting: src/core/matrix.cpp:70-96
Vec solve(Mat a, Vec b) {
  const std::size_t n = a.r;
  for (std::size_t col = 0; col < n; ++col) {
    std::size_t piv = col;
    for (std::size_t r = col + 1; r < n; ++r)
      if (std::fabs(a(r, col)) > std::fabs(a(piv, col))) piv = r;
    if (std::fabs(a(piv, col)) < 1e-300) throw Failure(CCR_ERR_RUNTIME, "solveLinearSystem: singular matrix");
    if (piv != col) {
      for (std::size_t c = 0; c < n; ++c) std::swap(a(col, c), a(piv, c));
      std::swap(b[col], b[piv]);
    }
    for (std::size_t r = col + 1; r < n; ++r) {
      const double f = a(r, col) / a(col, col);
      if (f == 0.0) continue;
      for (std::size_t c = col; c < n; ++c) a(r, c) -= f * a(col, c);
      b[r] -= f * b[col];
    }
  }
  Vec x(n);
  for (std::size_t i = n; i-- > 0;) {
    double s = b[i];
    for (std::size_t c = i + 1; c < n; ++c) s -= a(i, c) * x[c];
    x[i] = s / a(i, i);
  }
  return x;
}

// leastSquares by normal equations: src/core/regression.cpp:8-27
Vec leastSquares(const Mat& design, const Vec& y, double ridge) {
  const std::size_t n = design.r, k = design.c;
  Mat g(k, k, 0.0);
  Vec rhs(k, 0.0);
  for (std::size_t i = 0; i < n; ++i) {
    const double* row = &design.d[i * k];
    for (std::size_t a = 0; a < k; ++a) {
      rhs[a] += row[a] * y[i];
      for (std::size_t b = 0; b <= a; ++b) g(a, b) += row[a] * row[b];
    }
  }
  for (std::size_t a = 0; a < k; ++a) {
    for (std::size_t b = 0; b < a; ++b) g(b, a) = g(a, b);
    g(a, a) += ridge;
  }
  return solve(g, rhs);
}

// PolynomialRegression: src/core/regression.cpp:29-90 (standardise; degenerate regressor and
// small-sample degree fallback; ridge 1e-12 n; Horner predict).
struct PolyReg {
  std::size_t degree;
  double center = 0.0, scale = 1.0;
  Vec coef;
  explicit PolyReg(std::size_t d) : degree(d) {}
  void fit(const Vec& x, const Vec& y, const std::vector<char>* mask) {
    std::size_t n = 0;
    double sx = 0.0, sxx = 0.0, sy = 0.0;
    for (std::size_t i = 0; i < x.size(); ++i) {
      if (mask && !(*mask)[i]) continue;
      ++n;
      sx += x[i];
      sxx += x[i] * x[i];
      sy += y[i];
    }
    if (n == 0) {
      center = 0.0;
      scale = 1.0;
      coef = {0.0};
      return;
    }
    center = sx / static_cast<double>(n);
    const double var = std::max(sxx / static_cast<double>(n) - center * center, 0.0);
    scale = std::sqrt(var);
    std::size_t deg = degree;
    if (scale < 1e-12 * std::max(1.0, std::fabs(center))) deg = 0;
    if (n < 2 * (deg + 1)) deg = n >= 4 ? (n / 2) - 1 : 0;
    if (deg == 0) {
      scale = 1.0;
      coef = {sy / static_cast<double>(n)};
      return;
    }
    Mat design(n, deg + 1);
    Vec target(n);
    std::size_t row = 0;
    for (std::size_t i = 0; i < x.size(); ++i) {
      if (mask && !(*mask)[i]) continue;
      const double z = (x[i] - center) / scale;
      double pw = 1.0;
      for (std::size_t d = 0; d <= deg; ++d) {
        design(row, d) = pw;
        pw *= z;
      }
      target[row] = y[i];
      ++row;
    }
    coef = leastSquares(design, target, 1e-12 * static_cast<double>(n));
  }
  double predict(double x) const {
    const double z = (x - center) / scale;
    double r = 0.0;
    for (std::size_t d = coef.size(); d-- > 0;) r = r * z + coef[d];
    return r;
  }
};

// quantile: include/ccr/core/math.hpp:39-47 (sort, h = p (n - 1), linear interpolation).
double quantile(Vec v, double p) {
  if (v.empty()) invalid("quantile: empty sample");
  if (p < 0.0 || p > 1.0) invalid("quantile: level must be in [0, 1]");
  std::sort(v.begin(), v.end());
  const double h = p * static_cast<double>(v.size() - 1);
  const auto lo = static_cast<std::size_t>(std::floor(h));
  const std::size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (h - static_cast<double>(lo)) * (v[hi] - v[lo]);
}

// normalCdf and blackFormula: include/ccr/core/math.hpp:15-31. N(x) = erfc(-x / sqrt 2) / 2;
// discount times E[(s (F_T - K))^+] with total log-sd stdDev, intrinsic value when degenerate.
double normalCdf(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }
double blackFormula(bool call, double forward, double strike, double stdDev, double discount) {
  const double sign = call ? 1.0 : -1.0;
  if (stdDev <= 1e-14 || forward <= 0.0 || strike <= 0.0) return discount * std::max(sign * (forward - strike), 0.0);
  const double d1 = std::log(forward / strike) / stdDev + 0.5 * stdDev;
  const double d2 = d1 - stdDev;
  return discount * sign * (forward * normalCdf(sign * d1) - strike * normalCdf(sign * d2));
}

// ---------------------------------------------------------------- time grid
// src/core/time_grid.cpp:10-66
constexpr double kTimeTol = 1e-9;
struct Grid {
  Vec t{0.0};
  Grid() = default;
  explicit Grid(Vec times) {
    times.erase(std::remove_if(times.begin(), times.end(), [](double x) { return x < 0.0 || !std::isfinite(x); }),
                times.end());
    times.push_back(0.0);
    std::sort(times.begin(), times.end());
    t.clear();
    for (double x : times)
      if (t.empty() || x - t.back() > kTimeTol) t.push_back(x);
  }
  static Grid uniform(double horizon, std::size_t steps) {
    if (horizon <= 0.0 || steps == 0) invalid("TimeGrid::uniform: invalid arguments");
    Vec v(steps + 1);
    for (std::size_t i = 0; i <= steps; ++i) v[i] = horizon * static_cast<double>(i) / static_cast<double>(steps);
    return Grid(std::move(v));
  }
  Grid merged(const Vec& extra) const {
    Vec v = t;
    v.insert(v.end(), extra.begin(), extra.end());
    return Grid(std::move(v));
  }
  Grid lagged(double lag) const {
    if (lag <= 0.0) return *this;
    Vec v = t;
    for (double x : t)
      if (x > lag) v.push_back(x - lag);
    return Grid(std::move(v));
  }
  std::size_t size() const { return t.size(); }
  double operator[](std::size_t i) const { return t[i]; }
  double horizon() const { return t.back(); }
  std::size_t indexAtOrBefore(double x) const {
    const auto it = std::upper_bound(t.begin(), t.end(), x + kTimeTol);
    if (it == t.begin()) return 0;
    return static_cast<std::size_t>(it - t.begin()) - 1;
  }
  std::optional<std::size_t> find(double x, double tol = 1e-9) const {
    const std::size_t i = indexAtOrBefore(x);
    if (std::fabs(t[i] - x) <= tol) return i;
    if (i + 1 < t.size() && std::fabs(t[i + 1] - x) <= tol) return i + 1;
    return std::nullopt;
  }
};

// ---------------------------------------------------------------- market
// YieldCurve: src/market/yield_curve.cpp:10-45
struct Curve {
  Vec t, r;
  Curve() = default;
  explicit Curve(const ccr_yield_curve& c) {
    if (c.size <= 0 || !c.times || !c.zero_rates) invalid("YieldCurve: times and rates must be non-empty");
    t.assign(c.times, c.times + c.size);
    r.assign(c.zero_rates, c.zero_rates + c.size);
    for (std::size_t i = 1; i < t.size(); ++i)
      if (t[i] <= t[i - 1]) invalid("YieldCurve: times must be strictly increasing");
  }
  double zeroRate(double x) const {
    if (x <= t.front()) return r.front();
    if (x >= t.back()) return r.back();
    const std::size_t i = static_cast<std::size_t>(std::upper_bound(t.begin(), t.end(), x) - t.begin());
    const double w = (x - t[i - 1]) / (t[i] - t[i - 1]);
    return r[i - 1] + w * (r[i] - r[i - 1]);
  }
  double discount(double x) const { return std::exp(-zeroRate(x) * x); }
  double instantaneousForward(double x) const {
    double slope = 0.0;
    if (x >= t.front() && x < t.back()) {
      const std::size_t i = static_cast<std::size_t>(std::upper_bound(t.begin(), t.end(), x) - t.begin());
      slope = (r[i] - r[i - 1]) / (t[i] - t[i - 1]);
    }
    return zeroRate(x) + x * slope;
  }
};

// CreditCurve: src/market/credit_curve.cpp:10-41
struct Credit {
  Vec p, h;
  Credit() = default;
  explicit Credit(const ccr_credit_curve& c) {
    if (c.size <= 0 || !c.pillars || !c.hazards) invalid("CreditCurve: pillars and hazards must be non-empty");
    p.assign(c.pillars, c.pillars + c.size);
    h.assign(c.hazards, c.hazards + c.size);
    for (std::size_t i = 0; i < p.size(); ++i) {
      if (p[i] <= (i == 0 ? 0.0 : p[i - 1])) invalid("CreditCurve: pillars must be positive and increasing");
      if (h[i] < 0.0) invalid("CreditCurve: hazard rates must be non-negative");
    }
  }
  double survival(double x) const {
    if (x <= 0.0) return 1.0;
    double integral = 0.0, prev = 0.0;
    for (std::size_t i = 0; i < p.size(); ++i) {
      const bool last = i + 1 == p.size();
      const double end = last ? x : std::min(x, p[i]);
      integral += h[i] * (end - prev);
      if (end >= x) break;
      prev = end;
    }
    return std::exp(-integral);
  }
  double defaultProbability(double a, double b) const { return survival(a) - survival(b); }
};

// ---------------------------------------------------------------- models
// HullWhite1F (G1++): src/models/hull_white.cpp:12-68
struct HW {
  Curve curve;
  double a = 0.0, s = 0.0;
  HW() = default;
  explicit HW(const ccr_hull_white& m) : curve(m.curve), a(m.mean_reversion), s(m.volatility) {
    if (a <= 1e-8) invalid("HullWhite1F: mean reversion must be positive");
    if (s < 0.0) invalid("HullWhite1F: volatility must be non-negative");
  }
  double B(double t, double T) const { return (1.0 - std::exp(-a * (T - t))) / a; }
  double V(double t, double T) const {
    const double tau = T - t;
    const double e1 = std::exp(-a * tau);
    return s * s / (a * a) * (tau + 2.0 / a * e1 - 0.5 / a * e1 * e1 - 1.5 / a);
  }
  double zeroBond(double t, double T, double x) const {
    if (T <= t) return 1.0;
    const double ratio = curve.discount(T) / curve.discount(t);
    return ratio * std::exp(0.5 * (V(t, T) - V(0.0, T) + V(0.0, t)) - B(t, T) * x);
  }
  // Exact joint Gaussian step of (x, int x) plus int phi (hull_white.cpp:40-68). `drift` is the
  // constant drift of x under the simulation measure (the quanto adjustment of a foreign rate,
  // an extension); drift == 0 is the reference step exactly.
  void evolve(double t0, double t1, double x0, double z1, double z2, double drift, double& x1,
              double& integrated) const {
    const double dt = t1 - t0;
    if (dt <= 0.0) invalid("HullWhite1F::evolve: t1 must exceed t0");
    const double e = std::exp(-a * dt);
    const double s2 = s * s;
    const double varX = s2 * (1.0 - e * e) / (2.0 * a);
    const double varI = s2 / (a * a) * (dt - 2.0 * (1.0 - e) / a + (1.0 - e * e) / (2.0 * a));
    const double cov = s2 / (2.0 * a * a) * (1.0 - e) * (1.0 - e);
    const double sdX = std::sqrt(varX);
    x1 = x0 * e + sdX * z1;
    double c1 = 0.0, c2 = std::sqrt(std::max(varI, 0.0));
    if (sdX > 0.0) {
      c1 = cov / sdX;
      c2 = std::sqrt(std::max(varI - c1 * c1, 0.0));
    }
    double ix = x0 * (1.0 - e) / a + c1 * z1 + c2 * z2;
    if (drift != 0.0) {
      x1 += drift * (1.0 - e) / a;
      ix += drift * (dt - (1.0 - e) / a) / a;
    }
    const double g2 = dt - 2.0 / a * (std::exp(-a * t0) - std::exp(-a * t1)) +
                      0.5 / a * (std::exp(-2.0 * a * t0) - std::exp(-2.0 * a * t1));
    const double iphi = std::log(curve.discount(t0) / curve.discount(t1)) + s2 / (2.0 * a * a) * g2;
    integrated = ix + iphi;
  }
};

// LognormalAsset: src/models/lognormal_asset.cpp:9-22. With a foreign rate the carry integral is
// the simulated foreign integrated short rate (extension).
struct Asset {
  double spot = 1.0, vol = 0.0;
  Curve carry;
  int foreign = -1;
  explicit Asset(const ccr_lognormal_asset& m) : spot(m.spot), vol(m.volatility), foreign(m.foreign_rate) {
    if (spot <= 0.0) invalid("LognormalAsset: spot must be positive");
    if (vol < 0.0) invalid("LognormalAsset: volatility must be non-negative");
    if (foreign < 0) carry = Curve(m.carry);
  }
  double carryDiscount(double t, double T) const { return carry.discount(T) / carry.discount(t); }
  double evolveWithCarry(double t0, double t1, double s0, double idr, double icarry, double z) const {
    const double dt = t1 - t0;
    const double drift = idr - icarry - 0.5 * vol * vol * dt;
    return s0 * std::exp(drift + vol * std::sqrt(dt) * z);
  }
  double evolve(double t0, double t1, double s0, double idr, double z) const {
    return evolveWithCarry(t0, t1, s0, idr, std::log(carry.discount(t0) / carry.discount(t1)), z);
  }
};

// CirIntensity (CIR++): src/models/cir_intensity.cpp:9-66
struct Cir {
  Credit market;
  double kappa, theta, xi, y0;
  std::size_t sub;
  explicit Cir(const ccr_cir_intensity& m)
      : market(m.market), kappa(m.kappa), theta(m.theta), xi(m.xi), y0(m.y0), sub(static_cast<std::size_t>(m.substeps)) {
    if (kappa <= 0.0 || theta < 0.0 || xi < 0.0 || y0 < 0.0) invalid("CirIntensity: invalid parameters");
    if (m.substeps <= 0) invalid("CirIntensity: substeps must be positive");
  }
  double cirSurvival(double t) const {
    if (t <= 0.0) return 1.0;
    if (xi == 0.0) {
      const double integral = theta * t + (y0 - theta) * (1.0 - std::exp(-kappa * t)) / kappa;
      return std::exp(-integral);
    }
    const double h = std::sqrt(kappa * kappa + 2.0 * xi * xi);
    const double eh = std::exp(h * t) - 1.0;
    const double den = 2.0 * h + (kappa + h) * eh;
    const double logA = (2.0 * kappa * theta / (xi * xi)) * std::log(2.0 * h * std::exp(0.5 * (kappa + h) * t) / den);
    const double b = 2.0 * eh / den;
    return std::exp(logA - b * y0);
  }
  double integratedShift(double t) const { return std::log(cirSurvival(t) / market.survival(t)); }
  double survival(double t, double iy) const { return std::exp(-iy - integratedShift(t)); }
  void evolve(double dt, double z, const double* extra, double& y, double& iy) const {
    const double h = dt / static_cast<double>(sub);
    double remaining = std::sqrt(dt) * z;
    double remT = dt;
    for (std::size_t k = 0; k < sub; ++k) {
      double dW = remaining;
      if (k + 1 < sub) dW = h / remT * remaining + std::sqrt(h * (remT - h) / remT) * extra[k];
      remaining -= dW;
      remT -= h;
      const double yp = std::max(y, 0.0);
      const double yn = y + kappa * (theta - yp) * h + xi * std::sqrt(yp) * dW;
      iy += 0.5 * (yp + std::max(yn, 0.0)) * h;
      y = yn;
    }
  }
};

// ---------------------------------------------------------------- scenario generator
// Replayed normals (tests only): when set, ScenarioGenerator::generate consumes these in its
// draw order instead of the mt19937_64 stream.
thread_local const double* gReplay = nullptr;
thread_local std::size_t gReplayLen = 0;

// ScenarioGenerator: src/models/scenario_generator.cpp:22-143 (plus foreign rates).
struct Scenarios {
  Grid grid;
  std::size_t P = 0;
  Mat x, D;                                 // domestic state, deflator [path][date]
  std::vector<Mat> xf, S, y, Q;             // foreign states, assets, CIR y, survival
};

struct Generator {
  HW rates;
  std::vector<HW> foreign;
  std::vector<Asset> assets;
  std::vector<Cir> credits;
  Mat corr, L;
  std::vector<double> quanto;  // drift of each foreign state
  std::vector<int> fxOf;       // asset converting foreign currency f, -1 if none

  explicit Generator(const ccr_scenario_generator& g) : rates(g.rates) {
    if (g.num_foreign_rates < 0 || g.num_assets < 0 || g.num_credits < 0) invalid("ScenarioGenerator: negative count");
    for (int f = 0; f < g.num_foreign_rates; ++f) foreign.emplace_back(g.foreign_rates[f]);
    for (int a = 0; a < g.num_assets; ++a) {
      assets.emplace_back(g.assets[a]);
      if (assets.back().foreign >= static_cast<int>(foreign.size())) invalid("LognormalAsset: unknown foreign rate");
    }
    for (int c = 0; c < g.num_credits; ++c) credits.emplace_back(g.credits[c]);
    const std::size_t n = numFactors();
    corr = Mat(n, n, 0.0);
    if (g.correlation) {
      for (std::size_t i = 0; i < n * n; ++i) corr.d[i] = g.correlation[i];
    } else {
      for (std::size_t i = 0; i < n; ++i) corr(i, i) = 1.0;
    }
    for (std::size_t i = 0; i < n; ++i)
      if (std::fabs(corr(i, i) - 1.0) > 1e-12) invalid("ScenarioGenerator: correlation matrix must have unit diagonal");
    L = cholesky(corr);
    fxOf.assign(foreign.size(), -1);
    for (std::size_t a = 0; a < assets.size(); ++a)
      if (assets[a].foreign >= 0 && fxOf[assets[a].foreign] < 0) fxOf[assets[a].foreign] = static_cast<int>(a);
    quanto.assign(foreign.size(), 0.0);
    for (std::size_t f = 0; f < foreign.size(); ++f)
      if (fxOf[f] >= 0) {
        const std::size_t a = static_cast<std::size_t>(fxOf[f]);
        // Under the domestic measure the foreign state drifts by -rho(x_f, S_a) sigma_f sigma_S,
        // S_a the price of one unit of currency f in domestic units. The mutant build (test
        // infrastructure: a planted error the multi-currency martingale tests must catch)
        // flips the sign.
#ifdef CCR_ORACLE_MUTANT_QUANTO_SIGN
        quanto[f] = corr(1 + f, 1 + foreign.size() + a) * foreign[f].s * assets[a].vol;
#else
        quanto[f] = -corr(1 + f, 1 + foreign.size() + a) * foreign[f].s * assets[a].vol;
#endif
      }
  }
  std::size_t numFactors() const { return 1 + foreign.size() + assets.size() + credits.size(); }
  std::size_t numNormals() const {
    std::size_t n = numFactors() + 1;
    for (const auto& c : credits) n += c.sub - 1;
    return n + foreign.size();
  }

  Scenarios generate(const Grid& grid, std::size_t nP, std::uint64_t seed, bool antithetic, double* drawSeconds) const {
    if (nP == 0) invalid("ScenarioGenerator::generate: numPaths must be positive");
    const std::size_t nT = grid.size(), nF = foreign.size(), nA = assets.size(), nC = credits.size();
    const std::size_t nCorr = numFactors();
    std::vector<std::size_t> creditOff(nC);
    std::size_t nz = nCorr + 1;
    for (std::size_t c = 0; c < nC; ++c) {
      creditOff[c] = nz;
      nz += credits[c].sub - 1;
    }
    const std::size_t foreignOff = nz;
    nz += nF;

    Scenarios s;
    s.grid = grid;
    s.P = nP;
    s.x = Mat(nP, nT, 0.0);
    s.D = Mat(nP, nT, 1.0);
    s.xf.assign(nF, Mat(nP, nT, 0.0));
    s.S.assign(nA, Mat(nP, nT, 0.0));
    s.y.assign(nC, Mat(nP, nT, 0.0));
    s.Q.assign(nC, Mat(nP, nT, 1.0));

    std::mt19937_64 engine(seed);
    std::normal_distribution<double> normal(0.0, 1.0);
    // Draws come from the seeded stream, or from the replayed normals of this thread when a test
    // has installed some (each generate() starts at their beginning, as each one re-seeds).
    std::size_t replayPos = 0;
    auto draw = [&]() -> double {
      if (!gReplay) return normal(engine);
      if (replayPos >= gReplayLen) invalid("replayed normals exhausted");
      return gReplay[replayPos++];
    };
    const std::size_t steps = nT > 0 ? nT - 1 : 0;
    Vec z(steps * nz), w(nCorr), spot(nA), xf(nF), y(nC), iy(nC), irf(nF);
    double drawS = 0.0;
    for (std::size_t p = 0; p < nP; ++p) {
      const bool mirror = antithetic && (p % 2 == 1);
      if (mirror) {
        for (double& v : z) v = -v;
      } else {
        const auto t0 = std::chrono::steady_clock::now();
        for (double& v : z) v = draw();
        drawS += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
      double x = 0.0, defl = 1.0;
      for (std::size_t a = 0; a < nA; ++a) {
        spot[a] = assets[a].spot;
        s.S[a](p, 0) = spot[a];
      }
      for (std::size_t f = 0; f < nF; ++f) xf[f] = 0.0;
      for (std::size_t c = 0; c < nC; ++c) {
        y[c] = credits[c].y0;
        iy[c] = 0.0;
        s.y[c](p, 0) = y[c];
      }
      for (std::size_t j = 1; j < nT; ++j) {
        const double t0 = grid[j - 1], t1 = grid[j];
        const double* zz = z.data() + (j - 1) * nz;
        for (std::size_t i = 0; i < nCorr; ++i) {
          double sum = 0.0;
          for (std::size_t k = 0; k <= i; ++k) sum += L(i, k) * zz[k];
          w[i] = sum;
        }
        double ir = 0.0;
        rates.evolve(t0, t1, x, w[0], zz[nCorr], 0.0, x, ir);
        defl *= std::exp(-ir);
        s.x(p, j) = x;
        s.D(p, j) = defl;
        for (std::size_t f = 0; f < nF; ++f) {
          foreign[f].evolve(t0, t1, xf[f], w[1 + f], zz[foreignOff + f], quanto[f], xf[f], irf[f]);
          s.xf[f](p, j) = xf[f];
        }
        for (std::size_t a = 0; a < nA; ++a) {
          const Asset& as = assets[a];
          spot[a] = as.foreign < 0 ? as.evolve(t0, t1, spot[a], ir, w[1 + nF + a])
                                   : as.evolveWithCarry(t0, t1, spot[a], ir, irf[as.foreign], w[1 + nF + a]);
          s.S[a](p, j) = spot[a];
        }
        for (std::size_t c = 0; c < nC; ++c) {
          credits[c].evolve(t1 - t0, w[1 + nF + nA + c], zz + creditOff[c], y[c], iy[c]);
          s.y[c](p, j) = y[c];
          s.Q[c](p, j) = credits[c].survival(t1, iy[c]);
        }
      }
    }
    if (drawSeconds) *drawSeconds = drawS;
    return s;
  }
};

// ---------------------------------------------------------------- trades
constexpr double kEps = 1e-10;

struct Swap {
  double notional = 0.0, K = 0.0;
  Vec sched, fsched;  // fsched empty: single schedule (the reference swap)
  double sign = 1.0;
  int ccy = 0;        // 0 domestic, f + 1 foreign
  double maturity() const { return std::max(sched.back(), fsched.empty() ? sched.back() : fsched.back()); }
  Swap tail(double from) const {
    Swap s = *this;
    s.sched.clear();
    for (double t : sched)
      if (t >= from - kEps) s.sched.push_back(t);
    if (s.sched.size() < 2) invalid("InterestRateSwap::tail: no period starts after fromTime");
    s.fsched.clear();
    return s;
  }
};

void checkSchedule(const Vec& s, const char* what) {
  if (s.size() < 2) invalid(std::string(what) + ": schedule needs at least two dates");
  if (s.front() < 0.0) invalid(std::string(what) + ": schedule must start at t >= 0");
  for (std::size_t i = 1; i < s.size(); ++i)
    if (s[i] <= s[i - 1]) invalid(std::string(what) + ": schedule must be strictly increasing");
}

// InterestRateSwap::valueAtTime for one path: src/instruments/interest_rate_swap.cpp:47-78. The
// two-leg form (fsched) values the floating periods of fsched and the fixed coupons of sched with
// the same per-period rules.
double swapValue(const Swap& sw, const HW& m, const Scenarios& s, const Mat& X, std::size_t p, std::size_t j) {
  const double t = s.grid[j];
  double out = 0.0;
  if (t >= sw.maturity() - kEps) return 0.0;
  const double xj = X(p, j);
  if (sw.fsched.empty()) {
    const Vec& sc = sw.sched;
    for (std::size_t i = 1; i < sc.size(); ++i) {
      const double ts = sc[i - 1], te = sc[i];
      if (te <= t + kEps) continue;
      const double tau = te - ts;
      if (ts >= t - kEps) {
        const double pe = m.zeroBond(t, te, xj);
        const double floating = m.zeroBond(t, ts, xj) - pe;
        out += sw.sign * sw.notional * (floating - sw.K * tau * pe);
      } else {
        const std::size_t k = s.grid.indexAtOrBefore(ts);
        const double tk = s.grid[k];
        const double xk = X(p, k);
        const double fixing = (m.zeroBond(tk, ts, xk) / m.zeroBond(tk, te, xk) - 1.0) / tau;
        const double pe = m.zeroBond(t, te, xj);
        out += sw.sign * sw.notional * (fixing - sw.K) * tau * pe;
      }
    }
    return out;
  }
  const Vec& fl = sw.fsched;
  for (std::size_t i = 1; i < fl.size(); ++i) {
    const double ts = fl[i - 1], te = fl[i];
    if (te <= t + kEps) continue;
    const double tau = te - ts;
    if (ts >= t - kEps) {
      const double pe = m.zeroBond(t, te, xj);
      out += sw.sign * sw.notional * (m.zeroBond(t, ts, xj) - pe);
    } else {
      const std::size_t k = s.grid.indexAtOrBefore(ts);
      const double tk = s.grid[k];
      const double xk = X(p, k);
      const double fixing = (m.zeroBond(tk, ts, xk) / m.zeroBond(tk, te, xk) - 1.0) / tau;
      out += sw.sign * sw.notional * fixing * tau * m.zeroBond(t, te, xj);
    }
  }
  const Vec& fx = sw.sched;
  for (std::size_t i = 1; i < fx.size(); ++i) {
    const double ts = fx[i - 1], te = fx[i];
    if (te <= t + kEps) continue;
    out += sw.sign * sw.notional * (-sw.K * (te - ts) * m.zeroBond(t, te, xj));
  }
  return out;
}

struct Trade {
  int kind = CCR_TRADE_SWAP;
  Swap swap;
  // forward / European option (maturity = expiry)
  int asset = 0;
  double notional = 0.0, strike = 0.0, maturity = 0.0;
  int optionType = CCR_OPTION_CALL;
  // Bermudan
  Vec exercise;
  bool cash = false;
  double position = 1.0;
  std::size_t degree = 3;

  Vec eventTimes() const {
    if (kind == CCR_TRADE_ASSET_FORWARD || kind == CCR_TRADE_EUROPEAN_OPTION) return {maturity};
    Vec e = swap.sched;
    e.insert(e.end(), swap.fsched.begin(), swap.fsched.end());
    if (kind == CCR_TRADE_BERMUDAN_SWAPTION) e.insert(e.end(), exercise.begin(), exercise.end());
    return e;
  }
};

Trade makeTrade(const ccr_trade& d, const Generator& g) {
  Trade t;
  t.kind = d.kind;
  if (d.kind == CCR_TRADE_SWAP || d.kind == CCR_TRADE_BERMUDAN_SWAPTION) {
    if (!d.schedule || d.schedule_size < 2) invalid("InterestRateSwap: schedule needs at least two dates");
    t.swap.sched.assign(d.schedule, d.schedule + d.schedule_size);
    checkSchedule(t.swap.sched, "InterestRateSwap");
    if (d.float_schedule && d.float_schedule_size > 0) {
      if (d.kind != CCR_TRADE_SWAP) invalid("BermudanSwaption: the underlying has a single schedule");
      t.swap.fsched.assign(d.float_schedule, d.float_schedule + d.float_schedule_size);
      checkSchedule(t.swap.fsched, "InterestRateSwap");
    }
    t.swap.notional = d.notional;
    t.swap.K = d.fixed_rate;
    t.swap.sign = d.pay_fixed ? 1.0 : -1.0;
    t.swap.ccy = d.currency;
    if (d.currency < 0 || d.currency > static_cast<int>(g.foreign.size()))
      throw Failure(CCR_ERR_OUT_OF_RANGE, "InterestRateSwap: unknown currency");
    if (d.currency > 0 && g.fxOf[d.currency - 1] < 0)
      invalid("InterestRateSwap: no FX asset converts the swap currency");
  } else if (d.kind == CCR_TRADE_ASSET_FORWARD) {
    if (d.maturity <= 0.0) invalid("AssetForward: maturity must be positive");
    if (d.asset < 0 || d.asset >= static_cast<int>(g.assets.size()))
      throw Failure(CCR_ERR_OUT_OF_RANGE, "ScenarioSet: unknown asset");
    t.asset = d.asset;
    t.notional = d.notional;
    t.strike = d.fixed_rate;
    t.maturity = d.maturity;
  } else if (d.kind == CCR_TRADE_EUROPEAN_OPTION) {
    // EuropeanOption ctor: src/instruments/european_option.cpp:9-18
    if (d.maturity <= 0.0 || d.fixed_rate <= 0.0) invalid("EuropeanOption: invalid expiry or strike");
    if (d.asset < 0 || d.asset >= static_cast<int>(g.assets.size()))
      throw Failure(CCR_ERR_OUT_OF_RANGE, "ScenarioSet: unknown asset");
    if (g.assets[static_cast<std::size_t>(d.asset)].foreign >= 0)
      throw Failure(CCR_ERR_UNSUPPORTED, "EuropeanOption: the asset needs a deterministic carry curve");
    if (d.option_type != CCR_OPTION_CALL && d.option_type != CCR_OPTION_PUT) invalid("EuropeanOption: unknown type");
    t.asset = d.asset;
    t.notional = d.notional;
    t.strike = d.fixed_rate;
    t.maturity = d.maturity;
    t.optionType = d.option_type;
  } else {
    invalid("unknown trade kind");
  }
  if (d.kind == CCR_TRADE_BERMUDAN_SWAPTION) {
    // BermudanSwaption ctor: src/instruments/bermudan_swaption.cpp:17-32
    if (d.currency != 0) invalid("BermudanSwaption: domestic underlying only");
    if (!d.exercise_times || d.num_exercises <= 0) invalid("BermudanSwaption: no exercise dates");
    t.exercise.assign(d.exercise_times, d.exercise_times + d.num_exercises);
    std::sort(t.exercise.begin(), t.exercise.end());
    for (double e : t.exercise) {
      const bool on = std::any_of(t.swap.sched.begin(), t.swap.sched.end() - 1,
                                  [e](double x) { return std::fabs(x - e) < 1e-8; });
      if (!on) invalid("BermudanSwaption: exercise dates must be start dates of the underlying periods");
    }
    t.cash = d.cash_settled != 0;
    t.position = d.position;
    t.degree = static_cast<std::size_t>(std::max(0, d.regression_degree));
  }
  return t;
}

// BermudanSwaption::valueWithExercise: src/instruments/bermudan_swaption.cpp:49-146
struct BermudanValuation {
  Mat value;
  std::vector<long> exerciseIndex;
  double price = 0.0;
};

// Basis of the Bermudan regressions (ccr_oracle_set_regression_basis): -1 or 0 the reference's
// state basis (PolyReg); 1 the swap-rate / annuity basis below. Process-wide, set before a run.
int gBasis = -1;

// Swap-rate / annuity regression -- an extension with no reference counterpart, written on the
// rules of PolynomialRegression (src/core/regression.cpp:29-90) applied to two regressors:
// standardise each (mean, population sd over the sample); a degenerate one (sd at most 1e-6 of
// its root mean square) drops the fit to the sample mean; the total degree (the trade's degree, at most 3)
// falls while n < 2 x terms; design = the monomials of total degree <= d in (s, a) in the order
// 1, s, a, s^2, s a, a^2, s^3, s^2 a, s a^2, a^3; normal equations with the ridge 1e-12 n plus
// 1e-4 x their diagonal, solved by Cholesky.
struct SaReg {
  std::size_t degree;
  double m1 = 0.0, s1 = 1.0, m2 = 0.0, s2 = 1.0;
  std::size_t deg = 0;
  Vec coef;
  explicit SaReg(std::size_t d) : degree(d) {}
  static std::size_t terms(std::size_t d) { return (d + 1) * (d + 2) / 2; }
  static void row(double s, double a, double* phi) {
    const double s2 = s * s, a2 = a * a;
    phi[0] = 1.0;
    phi[1] = s;
    phi[2] = a;
    phi[3] = s2;
    phi[4] = s * a;
    phi[5] = a2;
    phi[6] = s2 * s;
    phi[7] = s2 * a;
    phi[8] = s * a2;
    phi[9] = a2 * a;
  }
  void fit(const Vec& u1, const Vec& u2, const Vec& y, const std::vector<char>* mask) {
    std::size_t n = 0;
    double a1 = 0.0, q1 = 0.0, a2 = 0.0, q2 = 0.0, sy = 0.0;
    for (std::size_t i = 0; i < u1.size(); ++i) {
      if (mask && !(*mask)[i]) continue;
      ++n;
      a1 += u1[i];
      q1 += u1[i] * u1[i];
      a2 += u2[i];
      q2 += u2[i] * u2[i];
      sy += y[i];
    }
    m1 = m2 = 0.0;
    s1 = s2 = 1.0;
    deg = 0;
    if (n == 0) {
      coef = {0.0};
      return;
    }
    const double nn = static_cast<double>(n);
    m1 = a1 / nn;
    m2 = a2 / nn;
    const double sd1 = std::sqrt(std::max(q1 / nn - m1 * m1, 0.0));
    const double sd2 = std::sqrt(std::max(q2 / nn - m2 * m2, 0.0));
    deg = std::min<std::size_t>(degree, 3);
    // Constant regressor: the one-pass variance of a constant carries rounding of ~1e-16 E[u^2];
    // a deviation of at most 1e-6 of the root mean square counts as none.
    if (!(sd1 > 1e-6 * std::sqrt(q1 / nn)) || !(sd2 > 1e-6 * std::sqrt(q2 / nn))) deg = 0;
    while (deg > 0 && n < 2 * terms(deg)) --deg;
    if (deg == 0) {
      coef = {sy / nn};
      return;
    }
    s1 = sd1;
    s2 = sd2;
    // Normal equations summed as leastSquares does (regression.cpp:8-27), the ridge 1e-12 n plus
    // 1e-4 x the diagonal (the monomials are numerically dependent when both regressors are
    // functions of one state), solved by Cholesky (matrix.cpp:49-68) and two triangular solves.
    const std::size_t k = terms(deg);
    Mat gm(k, k, 0.0);
    Vec rhs(k, 0.0);
    double phi[10];
    for (std::size_t i = 0; i < u1.size(); ++i) {
      if (mask && !(*mask)[i]) continue;
      row((u1[i] - m1) / s1, (u2[i] - m2) / s2, phi);
      for (std::size_t a = 0; a < k; ++a) {
        rhs[a] += phi[a] * y[i];
        for (std::size_t b = 0; b <= a; ++b) gm(a, b) += phi[a] * phi[b];
      }
    }
    for (std::size_t a = 0; a < k; ++a) {
      for (std::size_t b = 0; b < a; ++b) gm(b, a) = gm(a, b);
      gm(a, a) += 1e-12 * nn + 1e-4 * gm(a, a);
    }
    Mat l;
    try {
      l = cholesky(gm);
    } catch (const std::exception&) {
      deg = 0;
      s1 = s2 = 1.0;
      coef = {sy / nn};
      return;
    }
    Vec z(k);
    for (std::size_t i = 0; i < k; ++i) {
      double v = rhs[i];
      for (std::size_t m = 0; m < i; ++m) v -= l(i, m) * z[m];
      z[i] = v / l(i, i);
    }
    coef.assign(k, 0.0);
    for (std::size_t i = k; i-- > 0;) {
      double v = z[i];
      for (std::size_t m = i + 1; m < k; ++m) v -= l(m, i) * coef[m];
      coef[i] = v / l(i, i);
    }
  }
  double predict(double u1, double u2) const {
    if (deg == 0) return coef[0];
    double phi[10];
    row((u1 - m1) / s1, (u2 - m2) / s2, phi);
    double r = 0.0;
    for (std::size_t c = 0; c < coef.size(); ++c) r += coef[c] * phi[c];
    return r;
  }
};

// Swap rate F / A and annuity A at date j of path p of a tail swap: A = sum_i tau_i P(t, T_i) over
// its fixed periods, F = P(t, T_0) - P(t, T_n).
void swapRateAnnuityAt(const Vec& sched, const HW& hw, double t, double x, double& u1, double& u2) {
  double A = 0.0;
  for (std::size_t i = 1; i < sched.size(); ++i) A += (sched[i] - sched[i - 1]) * hw.zeroBond(t, sched[i], x);
  const double F = hw.zeroBond(t, sched.front(), x) - hw.zeroBond(t, sched.back(), x);
  u1 = F / A;
  u2 = A;
}
void swapRateAnnuity(const Swap& tail, const HW& hw, const Scenarios& s, std::size_t p, std::size_t j, double& u1,
                     double& u2) {
  swapRateAnnuityAt(tail.sched, hw, s.grid[j], s.x(p, j), u1, u2);
}

BermudanValuation bermudan(const Trade& tr, const Generator& g, const Scenarios& s) {
  const std::size_t nP = s.P, nT = s.grid.size(), nE = tr.exercise.size();
  std::vector<std::size_t> eIdx(nE);
  std::vector<Swap> tails;
  for (std::size_t k = 0; k < nE; ++k) {
    const auto idx = s.grid.find(tr.exercise[k], 1e-8);
    if (!idx) invalid("BermudanSwaption: exercise date not on the simulation grid");
    eIdx[k] = *idx;
    tails.push_back(tr.swap.tail(tr.exercise[k]));
  }
  std::vector<Vec> ev(nE, Vec(nP));
  for (std::size_t k = 0; k < nE; ++k)
    for (std::size_t p = 0; p < nP; ++p) ev[k][p] = swapValue(tails[k], g.rates, s, s.x, p, eIdx[k]);

  std::vector<Vec> strategy(nE + 1, Vec(nP, 0.0));
  std::vector<std::vector<char>> exercise(nE, std::vector<char>(nP, 0));
  Vec realised(nP, 0.0), target(nP), xcol(nP);
  std::vector<char> itm(nP);
  PolyReg reg(tr.degree);
  // Swap-rate / annuity basis: the regressors of the tail entered at exercise k, at date j.
  const bool sa = gBasis == 1;
  SaReg sreg(tr.degree);
  Vec u1(nP), u2(nP);
  auto features = [&](std::size_t k, std::size_t j) {
    for (std::size_t p = 0; p < nP; ++p) swapRateAnnuity(tails[k], g.rates, s, p, j, u1[p], u2[p]);
  };
  for (std::size_t k = nE; k-- > 0;) {
    const std::size_t j = eIdx[k];
    for (std::size_t p = 0; p < nP; ++p) {
      xcol[p] = s.x(p, j);
      target[p] = realised[p] / s.D(p, j);
      itm[p] = ev[k][p] > 0.0;
    }
    if (sa) {
      features(k, j);
      sreg.fit(u1, u2, target, &itm);
    } else {
      reg.fit(xcol, target, &itm);
    }
    for (std::size_t p = 0; p < nP; ++p) {
      const double e = ev[k][p];
      if (e > 0.0 && e > (sa ? sreg.predict(u1[p], u2[p]) : reg.predict(xcol[p]))) {
        exercise[k][p] = 1;
        realised[p] = s.D(p, j) * e;
      }
    }
    strategy[k] = realised;
  }
  BermudanValuation r;
  r.exerciseIndex.assign(nP, -1);
  for (std::size_t p = 0; p < nP; ++p)
    for (std::size_t k = 0; k < nE; ++k)
      if (exercise[k][p]) {
        r.exerciseIndex[p] = static_cast<long>(k);
        break;
      }
  r.value = Mat(nP, nT, 0.0);
  Vec cont(nP, 0.0);
  for (std::size_t j = 0; j < nT; ++j) {
    const double t = s.grid[j];
    std::size_t nextK = 0;
    while (nextK < nE && tr.exercise[nextK] <= t + kEps) ++nextK;
    if (nextK < nE) {
      for (std::size_t p = 0; p < nP; ++p) {
        xcol[p] = s.x(p, j);
        target[p] = strategy[nextK][p] / s.D(p, j);
      }
      if (sa) {
        features(nextK, j);
        sreg.fit(u1, u2, target, nullptr);
        for (std::size_t p = 0; p < nP; ++p) cont[p] = std::max(sreg.predict(u1[p], u2[p]), 0.0);
      } else {
        reg.fit(xcol, target, nullptr);
        for (std::size_t p = 0; p < nP; ++p) cont[p] = std::max(reg.predict(xcol[p]), 0.0);
      }
    } else {
      std::fill(cont.begin(), cont.end(), 0.0);
    }
    for (std::size_t p = 0; p < nP; ++p) {
      const long k = r.exerciseIndex[p];
      double v;
      if (k >= 0 && static_cast<std::size_t>(k) < nextK)
        v = tr.cash ? 0.0 : swapValue(tails[static_cast<std::size_t>(k)], g.rates, s, s.x, p, j);
      else
        v = cont[p];
      r.value(p, j) = tr.position * v;
    }
  }
  double sum = 0.0;
  for (std::size_t p = 0; p < nP; ++p) sum += r.value(p, 0);
  r.price = sum / static_cast<double>(nP);
  return r;
}

// Value of a pathwise trade on path p at date j (trade.cpp:5-13 with the per-trade valueAtTime).
double pathwiseValue(const Trade& tr, const Generator& g, const Scenarios& s, std::size_t p, std::size_t j) {
  if (tr.kind == CCR_TRADE_SWAP) {
    if (tr.swap.ccy == 0) return swapValue(tr.swap, g.rates, s, s.x, p, j);
    const std::size_t f = static_cast<std::size_t>(tr.swap.ccy - 1);
    const double vf = swapValue(tr.swap, g.foreign[f], s, s.xf[f], p, j);
    return vf * s.S[static_cast<std::size_t>(g.fxOf[f])](p, j);
  }
  const double t = s.grid[j];
  if (tr.kind == CCR_TRADE_EUROPEAN_OPTION) {
    // EuropeanOption::valueAtTime: src/instruments/european_option.cpp:20-35
    if (t >= tr.maturity - 1e-10) return 0.0;
    const Asset& as = g.assets[static_cast<std::size_t>(tr.asset)];
    const double carry = as.carryDiscount(t, tr.maturity);
    const double stdDev = as.vol * std::sqrt(tr.maturity - t);
    const double df = g.rates.zeroBond(t, tr.maturity, s.x(p, j));
    const double forward = s.S[static_cast<std::size_t>(tr.asset)](p, j) * carry / df;
    return tr.notional * blackFormula(tr.optionType == CCR_OPTION_CALL, forward, tr.strike, stdDev, df);
  }
  // AssetForward::valueAtTime: src/instruments/asset_forward.cpp:13-25
  if (t >= tr.maturity - 1e-10) return 0.0;
  const Asset& as = g.assets[static_cast<std::size_t>(tr.asset)];
  const double carry = as.foreign < 0 ? as.carryDiscount(t, tr.maturity)
                                      : g.foreign[static_cast<std::size_t>(as.foreign)].zeroBond(
                                            t, tr.maturity, s.xf[static_cast<std::size_t>(as.foreign)](p, j));
  const double spot = s.S[static_cast<std::size_t>(tr.asset)](p, j);
  return tr.notional * (spot * carry - tr.strike * g.rates.zeroBond(t, tr.maturity, s.x(p, j)));
}

// ---------------------------------------------------------------- collateral
// simulateCollateral: src/exposure/collateral.cpp:10-47
Mat collateral(const ccr_collateral& csa, const Mat& v, const Grid& grid, const Mat* price) {
  if (v.c != grid.size()) invalid("simulateCollateral: value matrix does not match the grid");
  if (csa.threshold_counterparty < 0.0 || csa.threshold_own < 0.0 || csa.minimum_transfer_amount < 0.0 ||
      csa.margin_period_of_risk < 0.0 || csa.haircut < 0.0 || csa.haircut >= 1.0)
    invalid("simulateCollateral: invalid collateral agreement");
  const bool sec = csa.type != CCR_COLLATERAL_CASH;
  if (sec && (!price || price->r != v.r || price->c != grid.size()))
    invalid("simulateCollateral: securities collateral needs the unit price on every path and date");
  const std::size_t nP = v.r, nT = grid.size();
  std::vector<std::size_t> call(nT);
  for (std::size_t j = 0; j < nT; ++j) call[j] = grid.indexAtOrBefore(std::max(grid[j] - csa.margin_period_of_risk, 0.0));
  auto required = [&csa](double value) {
    double c = 0.0;
    if (value > csa.threshold_counterparty) c += value - csa.threshold_counterparty;
    if (std::isfinite(csa.threshold_own) && -value > csa.threshold_own) c -= -value - csa.threshold_own;
    return c;
  };
  Mat out(nP, nT, 0.0);
  const double keep = sec ? 1.0 - csa.haircut : 1.0;
  auto pr = [&](std::size_t p, std::size_t j) { return sec ? (*price)(p, j) : 1.0; };
  for (std::size_t p = 0; p < nP; ++p) {
    double units = required(v(p, call[0])) / (keep * pr(p, call[0]));
    for (std::size_t j = 0; j < nT; ++j) {
      const std::size_t c = call[j];
      const double target = required(v(p, c));
      if (std::fabs(target - units * keep * pr(p, c)) >= csa.minimum_transfer_amount) units = target / (keep * pr(p, c));
      out(p, j) = units * pr(p, j) + csa.independent_amount;
    }
  }
  return out;
}

// ---------------------------------------------------------------- engine
Grid reportingGrid(const Grid& base, const std::vector<Trade>& trades) {
  Vec ev;
  for (const auto& t : trades) {
    const Vec e = t.eventTimes();
    ev.insert(ev.end(), e.begin(), e.end());
  }
  return base.merged(ev);
}

Grid simulationGrid(const Grid& base, const std::vector<Trade>& trades, const ccr_netting_set& ns) {
  Grid g = reportingGrid(base, trades);
  if (ns.has_collateral) g = g.lagged(ns.collateral.margin_period_of_risk);
  return g;
}

template <class F>
void parallelFor(std::size_t n, int threads, F&& body) {
  if (threads <= 1 || n < 2) {
    for (std::size_t i = 0; i < n; ++i) body(i);
    return;
  }
  const std::size_t nt = std::min<std::size_t>(static_cast<std::size_t>(threads), n);
  std::vector<std::thread> pool;
  for (std::size_t w = 0; w < nt; ++w)
    pool.emplace_back([&, w] {
      const std::size_t b = n * w / nt, e = n * (w + 1) / nt;
      for (std::size_t i = b; i < e; ++i) body(i);
    });
  for (auto& th : pool) th.join();
}

thread_local ccr_oracle_timing gTiming{};

double secondsSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ExposureEngine::run: src/exposure/exposure_engine.cpp:50-81 (plus the fused pipeline's per-date
// extras: E[D], E[Q], pathwise CVA/DVA increments on the reporting dates).
// im (optional): initialMarginProfile (src/xva/xva.cpp:59-76) of the netted value on the
// simulation grid, imQ its quantile level, imMpor the margin period of risk.
// tstats (optional): per trade i, kTradeStatCount rows of one value per reporting date
// ([i][m][r] in tstats[i][m * nR + r]): the statistics of gpu per-trade mode
// (src/gpu/fused_kernels.cpp:174-183) -- E[D V_i 1{V>0}], E[V_i 1{V>0}], E[D (V+ - (V - V_i)+)],
// E[D V_i+], E[V_i+] with V the netted value before collateral, the estimators of
// marginalExposureContributions and incrementalExposure (src/exposure/allocation.cpp:8-45).
constexpr int kTradeStatCount = 5;
void runSet(const Generator& g, const ccr_simulation_config& cfg, const Grid& base, const ccr_netting_set& ns,
            double q, int threads, ccr_exposure_out* out, Vec* im = nullptr, double imQ = 0.99,
            double imMpor = 10.0 / 250.0, std::vector<Vec>* tstats = nullptr, const Grid* repGrid = nullptr,
            const Grid* simGrid = nullptr) {
  if (ns.num_trades <= 0 || !ns.trades) invalid("ExposureEngine::run: empty netting set");
  std::vector<Trade> trades;
  for (int i = 0; i < ns.num_trades; ++i) trades.push_back(makeTrade(ns.trades[i], g));
  const int nC = static_cast<int>(g.credits.size());
  if (ns.counterparty_credit >= nC || ns.own_credit >= nC) throw Failure(CCR_ERR_OUT_OF_RANGE, "credit index out of range");
  const Grid rep = repGrid ? *repGrid : reportingGrid(base, trades);
  const Grid sim = simGrid ? *simGrid : simulationGrid(base, trades, ns);
  auto t0 = std::chrono::steady_clock::now();
  double drawS = 0.0;
  const Scenarios s = g.generate(sim, static_cast<std::size_t>(cfg.num_paths), cfg.seed, cfg.antithetic != 0, &drawS);
  gTiming.draws_s += drawS;
  gTiming.simulate_s += secondsSince(t0);
  std::vector<std::size_t> ri;
  for (double t : rep.t) ri.push_back(*s.grid.find(t));
  const std::size_t nP = s.P, nT = s.grid.size(), nR = ri.size();

  t0 = std::chrono::steady_clock::now();
  // Bermudan cubes first (they need every path), then the netted value in trade order.
  std::vector<std::unique_ptr<BermudanValuation>> berm(trades.size());
  for (std::size_t i = 0; i < trades.size(); ++i)
    if (trades[i].kind == CCR_TRADE_BERMUDAN_SWAPTION) berm[i] = std::make_unique<BermudanValuation>(bermudan(trades[i], g, s));
  Mat net(nP, nT, 0.0);
  std::vector<Mat> tv(tstats ? trades.size() : 0, Mat(tstats ? nP : 0, tstats ? nT : 0, 0.0));
  // Under a CSA every simulation date is needed (margin-call look-back); otherwise the
  // simulation grid is the reporting grid.
  parallelFor(nP, threads, [&](std::size_t p) {
    for (std::size_t j = 0; j < nT; ++j) {
      double total = 0.0;
      for (std::size_t i = 0; i < trades.size(); ++i) {
        const double v = berm[i] ? berm[i]->value(p, j) : pathwiseValue(trades[i], g, s, p, j);
        if (tstats) tv[i](p, j) = v;
        total = i == 0 ? v : total + v;
      }
      net(p, j) = total;
    }
  });
  if (tstats) {
    tstats->assign(trades.size(), Vec(static_cast<std::size_t>(kTradeStatCount) * nR, 0.0));
    for (std::size_t i = 0; i < trades.size(); ++i)
      for (std::size_t r = 0; r < nR; ++r) {
        const std::size_t j = ri[r];
        double sum[kTradeStatCount] = {0, 0, 0, 0, 0};
        for (std::size_t p = 0; p < nP; ++p) {
          const double v = net(p, j), vi = tv[i](p, j), d = s.D(p, j);
          const double on = v > 0.0 ? 1.0 : 0.0;
          sum[0] += d * vi * on;
          sum[1] += vi * on;
          sum[2] += d * (std::max(v, 0.0) - std::max(v - vi, 0.0));
          sum[3] += d * std::max(vi, 0.0);
          sum[4] += std::max(vi, 0.0);
        }
        for (int m = 0; m < kTradeStatCount; ++m)
          (*tstats)[i][static_cast<std::size_t>(m) * nR + r] = sum[m] / static_cast<double>(nP);
      }
  }
  gTiming.pricing_s += secondsSince(t0);
  gTiming.revaluations += static_cast<int64_t>(trades.size()) * static_cast<int64_t>(nP) * static_cast<int64_t>(nR);
  if (im) {
    im->assign(nT, 0.0);
    for (std::size_t k = 1; k < nT; ++k) {
      Vec change(nP);
      double meanDeflator = 0.0;
      for (std::size_t p = 0; p < nP; ++p) {
        change[p] = std::fabs(net(p, k) - net(p, k - 1));
        meanDeflator += s.D(p, k);
      }
      const double scale = std::sqrt(imMpor / (s.grid[k] - s.grid[k - 1]));
      (*im)[k] = quantile(std::move(change), imQ) * scale * meanDeflator / static_cast<double>(nP);
    }
  }

  t0 = std::chrono::steady_clock::now();
  Mat coll(nP, nT, 0.0);
  if (ns.has_collateral) {
    const ccr_collateral& csa = ns.collateral;
    Mat price;
    if (csa.type == CCR_COLLATERAL_BOND) {
      if (csa.bond_maturity < s.grid.horizon() - 1e-10) invalid("the collateral bond must mature after the horizon");
      price = Mat(nP, nT);
      for (std::size_t p = 0; p < nP; ++p)
        for (std::size_t j = 0; j < nT; ++j) price(p, j) = g.rates.zeroBond(s.grid[j], csa.bond_maturity, s.x(p, j));
    } else if (csa.type == CCR_COLLATERAL_ASSET) {
      if (csa.asset_index < 0 || csa.asset_index >= static_cast<int>(g.assets.size()))
        throw Failure(CCR_ERR_OUT_OF_RANGE, "collateral asset index out of range");
      price = s.S[static_cast<std::size_t>(csa.asset_index)];
    }
    coll = collateral(csa, net, s.grid, csa.type == CCR_COLLATERAL_CASH ? nullptr : &price);
  }
  gTiming.collateral_s += secondsSince(t0);

  t0 = std::chrono::steady_clock::now();
  // ExposureProfile::compute on the reporting columns: src/exposure/exposure_profile.cpp:10-49
  if (out->capacity < static_cast<int32_t>(nR)) invalid("output capacity smaller than the reporting grid");
  out->num_dates = static_cast<int32_t>(nR);
  const double n = static_cast<double>(nP);
  Vec pos(nP);
  const Mat* Qc = ns.counterparty_credit >= 0 ? &s.Q[static_cast<std::size_t>(ns.counterparty_credit)] : nullptr;
  const Mat* Qo = ns.own_credit >= 0 ? &s.Q[static_cast<std::size_t>(ns.own_credit)] : nullptr;
  for (std::size_t r = 0; r < nR; ++r) {
    const std::size_t j = ri[r];
    double sum = 0.0, ee = 0.0, ene = 0.0, dee = 0.0, dene = 0.0, md = 0.0, mq = 0.0, cva = 0.0, dva = 0.0;
    for (std::size_t p = 0; p < nP; ++p) {
      const double v = net(p, j) - coll(p, j);
      const double d = s.D(p, j);
      sum += v;
      pos[p] = std::max(v, 0.0);
      ee += pos[p];
      ene += std::min(v, 0.0);
      dee += d * pos[p];
      dene += d * std::min(v, 0.0);
      md += d;
      if (Qc) mq += (*Qc)(p, j);
      if (r > 0) {
        const std::size_t i0 = ri[r - 1];
        const double v0 = net(p, i0) - coll(p, i0);
        if (Qc) {
          const double prev = s.D(p, i0) * std::max(v0, 0.0), cur = d * pos[p];
          cva += 0.5 * (prev + cur) * ((*Qc)(p, i0) - (*Qc)(p, j));
        }
        if (Qo) {
          const double prev = s.D(p, i0) * std::max(-v0, 0.0), cur = d * std::max(-v, 0.0);
          dva += 0.5 * (prev + cur) * ((*Qo)(p, i0) - (*Qo)(p, j));
        }
      }
      if (out->exposure_rows) out->exposure_rows[r * nP + p] = pos[p];
    }
    if (out->times) out->times[r] = rep.t[r];
    if (out->expected_value) out->expected_value[r] = sum / n;
    if (out->expected_exposure) out->expected_exposure[r] = ee / n;
    if (out->expected_negative_exposure) out->expected_negative_exposure[r] = ene / n;
    if (out->discounted_expected_exposure) out->discounted_expected_exposure[r] = dee / n;
    if (out->discounted_expected_negative_exposure) out->discounted_expected_negative_exposure[r] = dene / n;
    if (out->mean_deflator) out->mean_deflator[r] = md / n;
    if (out->mean_survival) out->mean_survival[r] = Qc ? mq / n : 1.0;
    if (out->cva_increments) out->cva_increments[r] = cva / n;
    if (out->dva_increments) out->dva_increments[r] = dva / n;
    if (out->potential_future_exposure) out->potential_future_exposure[r] = quantile(pos, q);
  }
  gTiming.statistics_s += secondsSince(t0);
}

int32_t guard(char* err, int32_t cap, const std::function<void()>& f) {
  try {
    f();
    if (err && cap > 0) err[0] = 0;
    return CCR_OK;
  } catch (const Failure& e) {
    if (err && cap > 0) std::snprintf(err, static_cast<std::size_t>(cap), "%s", e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (err && cap > 0) std::snprintf(err, static_cast<std::size_t>(cap), "%s", e.what());
    return CCR_ERR_RUNTIME;
  }
}

int32_t writeGrid(const Grid& g, double* out, int32_t cap, int32_t* size) {
  if (size) *size = static_cast<int32_t>(g.size());
  if (out)
    for (std::size_t i = 0; i < g.size() && static_cast<int32_t>(i) < cap; ++i) out[i] = g[i];
  return CCR_OK;
}

std::vector<Trade> tradesOf(const ccr_netting_set* set, const Generator* g) {
  std::vector<Trade> trades;
  for (int i = 0; i < set->num_trades; ++i) {
    Trade t;
    const ccr_trade& d = set->trades[i];
    t.kind = d.kind;
    if (d.kind == CCR_TRADE_ASSET_FORWARD || d.kind == CCR_TRADE_EUROPEAN_OPTION) {
      t.maturity = d.maturity;
    } else {
      if (d.schedule && d.schedule_size > 0) t.swap.sched.assign(d.schedule, d.schedule + d.schedule_size);
      if (d.float_schedule && d.float_schedule_size > 0)
        t.swap.fsched.assign(d.float_schedule, d.float_schedule + d.float_schedule_size);
      if (d.exercise_times && d.num_exercises > 0) t.exercise.assign(d.exercise_times, d.exercise_times + d.num_exercises);
    }
    trades.push_back(t);
  }
  (void)g;
  return trades;
}

// ---------------------------------------------------------------- nested initial margin
// The nested conditional initial margin of the fused pipeline (gpu::FusedOptions::imInnerPaths,
// src/gpu/fused_kernels.cpp:445-479, the CPU mirror src/gpu/fused_reference.cpp:371-402; the grid
// extension src/gpu/fused_exposure.cpp:199-212) restated in FP64 on the engine's scenarios: from
// every path and reporting date t whose t + mpor lies on the simulation grid, `inner` inner paths
// (antithetic pairs) continue the path's rates and assets over (t, t + mpor] and revalue the
// netting set there (fixings inside the window from the inner path, earlier ones from the outer
// path); IM = the interpolated order statistic of |V_inner(t + mpor) - V(t)| at level q, read from
// the 16 largest samples as the kernel does (the top-16 rule of gpu::compile guarantees the
// statistic is among them). The inner normals are this build's own counter-based generator (the
// reference's fused pipeline keys PCG draws by its own stream layout; an FP64 restatement of the
// engine cannot share them): SplitMix64's finaliser chained over (seed, set, path, date, pair),
// normal k of inner step j by the Box-Muller cosine branch of two uniforms.
std::uint64_t nestedMix(std::uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double nestedNormal(std::uint64_t key, std::size_t step, std::size_t k) {
  const std::uint64_t c = key ^ ((static_cast<std::uint64_t>(step) << 16) | static_cast<std::uint64_t>(k));
  const std::uint64_t a = nestedMix(c), b = nestedMix(c ^ 0xD1B54A32D192ED03ull);
  const double u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// The simulation grid of a netting set with the nested initial margin: the engine's grid merged
// with t + mpor for every reporting date t up to its horizon (fused_exposure.cpp:199-212).
Grid nestedSimulationGrid(const Grid& rep, const Grid& sim, double mpor) {
  Vec ahead;
  for (double t : rep.t)
    if (t + mpor <= sim.t.back() + 1e-10) ahead.push_back(t + mpor);
  Vec v = sim.t;
  v.insert(v.end(), ahead.begin(), ahead.end());
  return Grid(v);
}

struct NestedResult {
  Vec discounted, mean, b05, b50, b95;
  Vec samples;  // the inner samples |V_inner - V| of (samplePath, sampleRep), when asked
  double sampleIm = 0.0;
};

NestedResult nestedIm(const Generator& g, const ccr_simulation_config& cfg, const Grid& base, const ccr_netting_set& ns,
                      int setIndex, int inner, double q, double mpor, int threads, long samplePath, int sampleRep) {
  if (ns.num_trades <= 0 || !ns.trades) invalid("ExposureEngine::run: empty netting set");
  std::vector<Trade> trades;
  for (int i = 0; i < ns.num_trades; ++i) trades.push_back(makeTrade(ns.trades[i], g));
  for (const auto& t : trades)
    if (t.kind == CCR_TRADE_BERMUDAN_SWAPTION) throw Failure(CCR_ERR_UNSUPPORTED, "nested initial margin: Bermudan swaptions");
  const Grid rep = reportingGrid(base, trades);
  const Grid sim = nestedSimulationGrid(rep, simulationGrid(base, trades, ns), mpor);
  const Scenarios s = g.generate(sim, static_cast<std::size_t>(cfg.num_paths), cfg.seed, cfg.antithetic != 0, nullptr);
  const std::size_t nP = s.P, nT = s.grid.size(), nR = rep.size();
  const std::size_t nF = g.foreign.size(), nA = g.assets.size(), nCorr = g.numFactors();
  std::size_t nz = nCorr + 1;
  for (const auto& c : g.credits) nz += c.sub - 1;
  const std::size_t foreignOff = nz;
  nz += nF;
  std::vector<std::size_t> ri(nR), target(nR, nT);
  for (std::size_t r = 0; r < nR; ++r) {
    ri[r] = *s.grid.find(rep[r]);
    if (const auto k = s.grid.find(s.grid[ri[r]] + mpor, 1e-8)) target[r] = *k;
  }
  auto value = [&](const Scenarios& sc, std::size_t p, std::size_t j) {
    double total = 0.0;
    for (std::size_t i = 0; i < trades.size(); ++i) {
      const double v = pathwiseValue(trades[i], g, sc, p, j);
      total = i == 0 ? v : total + v;
    }
    return total;
  };
  Mat im(nP, nR, 0.0);
  NestedResult res;
  const double rank = q * static_cast<double>(inner - 1);
  const auto kq = static_cast<std::size_t>(std::floor(rank));
  const std::size_t hiIdx = std::min<std::size_t>(static_cast<std::size_t>(inner) - 1 - kq, 15);
  std::mutex sampleLock;
  parallelFor(nP, threads, [&](std::size_t p) {
    // One path's scenario (P = 1) holding the outer path; an inner path overwrites (t, t + mpor].
    Scenarios w;
    w.grid = s.grid;
    w.P = 1;
    w.x = Mat(1, nT, 0.0);
    w.D = Mat(1, nT, 1.0);
    w.xf.assign(nF, Mat(1, nT, 0.0));
    w.S.assign(nA, Mat(1, nT, 0.0));
    for (std::size_t j = 0; j < nT; ++j) {
      w.x(0, j) = s.x(p, j);
      for (std::size_t f = 0; f < nF; ++f) w.xf[f](0, j) = s.xf[f](p, j);
      for (std::size_t a = 0; a < nA; ++a) w.S[a](0, j) = s.S[a](p, j);
    }
    const std::uint64_t pathKey = nestedMix(nestedMix(nestedMix(cfg.seed) ^ static_cast<std::uint64_t>(setIndex)) ^
                                            static_cast<std::uint64_t>(p));
    Vec zz(nz, 0.0), wv(nCorr), irf(nF), samples;
    for (std::size_t r = 0; r < nR; ++r) {
      const std::size_t j = ri[r], tg = target[r];
      if (tg >= nT) continue;  // t + mpor past the horizon: IM 0
      const double V = value(s, p, j);
      double top[16] = {};
      samples.clear();
      const std::uint64_t dateKey = nestedMix(pathKey ^ (static_cast<std::uint64_t>(j) << 32));
      for (int i = 0; i < inner; ++i) {
        const std::uint64_t key = nestedMix(dateKey ^ static_cast<std::uint64_t>(i >> 1));
        const double sign = (i & 1) ? -1.0 : 1.0;
        double x = s.x(p, j);
        Vec xf(nF), spot(nA);
        for (std::size_t f = 0; f < nF; ++f) xf[f] = s.xf[f](p, j);
        for (std::size_t a = 0; a < nA; ++a) spot[a] = s.S[a](p, j);
        for (std::size_t jj = j + 1; jj <= tg; ++jj) {
          // The step of ScenarioGenerator::generate (scenario_generator.cpp:118-134) for the rates
          // and assets on the inner normals (drivers, then the integral normals).
          for (std::size_t k = 0; k < nCorr; ++k) zz[k] = sign * nestedNormal(key, jj, k);
          zz[nCorr] = sign * nestedNormal(key, jj, nCorr);
          for (std::size_t f = 0; f < nF; ++f) zz[foreignOff + f] = sign * nestedNormal(key, jj, foreignOff + f);
          const double t0 = s.grid[jj - 1], t1 = s.grid[jj];
          for (std::size_t a = 0; a < nCorr; ++a) {
            double sum = 0.0;
            for (std::size_t k = 0; k <= a; ++k) sum += g.L(a, k) * zz[k];
            wv[a] = sum;
          }
          double ir = 0.0;
          g.rates.evolve(t0, t1, x, wv[0], zz[nCorr], 0.0, x, ir);
          w.x(0, jj) = x;
          for (std::size_t f = 0; f < nF; ++f) {
            g.foreign[f].evolve(t0, t1, xf[f], wv[1 + f], zz[foreignOff + f], g.quanto[f], xf[f], irf[f]);
            w.xf[f](0, jj) = xf[f];
          }
          for (std::size_t a = 0; a < nA; ++a) {
            const Asset& as = g.assets[a];
            spot[a] = as.foreign < 0 ? as.evolve(t0, t1, spot[a], ir, wv[1 + nF + a])
                                     : as.evolveWithCarry(t0, t1, spot[a], ir, irf[as.foreign], wv[1 + nF + a]);
            w.S[a](0, jj) = spot[a];
          }
        }
        double cur = std::fabs(value(w, 0, tg) - V);
        if (static_cast<long>(p) == samplePath && static_cast<int>(r) == sampleRep) samples.push_back(cur);
        for (int k = 0; k < 16; ++k)
          if (cur > top[k]) std::swap(cur, top[k]);
      }
      // Restore the outer path over the window.
      for (std::size_t jj = j + 1; jj <= tg; ++jj) {
        w.x(0, jj) = s.x(p, jj);
        for (std::size_t f = 0; f < nF; ++f) w.xf[f](0, jj) = s.xf[f](p, jj);
        for (std::size_t a = 0; a < nA; ++a) w.S[a](0, jj) = s.S[a](p, jj);
      }
      const double ok = top[hiIdx], ok1 = hiIdx > 0 ? top[hiIdx - 1] : ok;
      im(p, r) = ok + (rank - static_cast<double>(kq)) * (ok1 - ok);
      if (static_cast<long>(p) == samplePath && static_cast<int>(r) == sampleRep) {
        std::lock_guard<std::mutex> lk(sampleLock);
        res.samples = samples;
        res.sampleIm = im(p, r);
      }
    }
  });
  const double n = static_cast<double>(nP);
  for (Vec* v : {&res.discounted, &res.mean, &res.b05, &res.b50, &res.b95}) v->assign(nR, 0.0);
  for (std::size_t r = 0; r < nR; ++r) {
    double sd = 0.0, sm = 0.0;
    Vec col(nP);
    for (std::size_t p = 0; p < nP; ++p) {
      sd += s.D(p, ri[r]) * im(p, r) / n;
      sm += im(p, r) / n;
      col[p] = im(p, r);
    }
    res.discounted[r] = sd;
    res.mean[r] = sm;
    res.b05[r] = quantile(col, 0.05);
    res.b50[r] = quantile(col, 0.5);
    res.b95[r] = quantile(col, 0.95);
  }
  return res;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

int32_t ccr_oracle_run_exposure(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                const double* base_grid, int32_t base_size, const ccr_netting_set* sets,
                                int32_t num_sets, double pfe_quantile, int32_t threads, ccr_exposure_out* outs,
                                char* err, int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !sets || !outs || num_sets <= 0) invalid("null argument");
    gTiming = ccr_oracle_timing{};
    const auto t0 = std::chrono::steady_clock::now();
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    for (int i = 0; i < num_sets; ++i) runSet(g, *config, base, sets[i], pfe_quantile, threads, &outs[i]);
    gTiming.total_s = secondsSince(t0);
  });
}

// ccr_oracle_run_exposure on replayed normals instead of the seeded stream (tests of model
// reductions that need identical draws under two draw layouts).
int32_t ccr_oracle_run_exposure_replay(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                       const double* base_grid, int32_t base_size, const ccr_netting_set* sets,
                                       int32_t num_sets, double pfe_quantile, int32_t threads, const double* normals,
                                       int64_t num_normals, ccr_exposure_out* outs, char* err, int32_t err_capacity) {
  if (!normals || num_normals <= 0) {
    if (err && err_capacity > 0) std::snprintf(err, static_cast<std::size_t>(err_capacity), "no normals");
    return CCR_ERR_INVALID_ARGUMENT;
  }
  gReplay = normals;
  gReplayLen = static_cast<std::size_t>(num_normals);
  const int32_t rc = ccr_oracle_run_exposure(generator, config, base_grid, base_size, sets, num_sets, pfe_quantile,
                                             threads, outs, err, err_capacity);
  gReplay = nullptr;
  gReplayLen = 0;
  return rc;
}

// The scenario set of ScenarioGenerator::generate (src/models/scenario_generator.cpp:61-143) on
// `grid`, from the seeded stream or from replayed normals (normals NULL: the stream). Every
// output is [path][date] row-major and may be NULL; foreign states, spots and survivals are
// stacked per factor ([factor][path][date]).
int32_t ccr_oracle_scenarios(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* grid, int32_t grid_size, const double* normals, int64_t num_normals,
                             double* x, double* deflator, double* foreign_x, double* spots, double* survival,
                             char* err, int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !grid || grid_size <= 0) invalid("null argument");
    const Generator g(*generator);
    const Grid gr(Vec(grid, grid + grid_size));
    gReplay = normals;
    gReplayLen = normals ? static_cast<std::size_t>(num_normals) : 0;
    Scenarios sc;
    try {
      sc = g.generate(gr, static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0, nullptr);
    } catch (...) {
      gReplay = nullptr;
      gReplayLen = 0;
      throw;
    }
    gReplay = nullptr;
    gReplayLen = 0;
    const std::size_t n = sc.P * gr.size();
    auto put = [&](double* dst, const std::vector<Mat>& ms) {
      if (!dst) return;
      for (std::size_t f = 0; f < ms.size(); ++f) std::copy(ms[f].d.begin(), ms[f].d.end(), dst + f * n);
    };
    if (x) std::copy(sc.x.d.begin(), sc.x.d.end(), x);
    if (deflator) std::copy(sc.D.d.begin(), sc.D.d.end(), deflator);
    put(foreign_x, sc.xf);
    put(spots, sc.S);
    put(survival, sc.Q);
  });
}

// Portfolio book (src/app/analyses.cpp:1690-1709, 1727-1766): every set reports on the union of
// the sets' reporting grids and simulates the union of their simulation grids built on it.
int32_t ccr_oracle_run_book(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                            const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t num_sets,
                            double pfe_quantile, int32_t threads, ccr_exposure_out* outs, char* err,
                            int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !sets || !outs || num_sets <= 0) invalid("null argument");
    gTiming = ccr_oracle_timing{};
    const auto t0 = std::chrono::steady_clock::now();
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    std::vector<std::vector<Trade>> trades(static_cast<std::size_t>(num_sets));
    Vec r, v;
    for (int i = 0; i < num_sets; ++i) {
      for (int k = 0; k < sets[i].num_trades; ++k) trades[static_cast<std::size_t>(i)].push_back(makeTrade(sets[i].trades[k], g));
      const Grid gr = reportingGrid(base, trades[static_cast<std::size_t>(i)]);
      r.insert(r.end(), gr.t.begin(), gr.t.end());
    }
    const Grid rep(r);
    for (int i = 0; i < num_sets; ++i) {
      const Grid gs = simulationGrid(rep, trades[static_cast<std::size_t>(i)], sets[i]);
      v.insert(v.end(), gs.t.begin(), gs.t.end());
    }
    const Grid sim(v);
    for (int i = 0; i < num_sets; ++i)
      runSet(g, *config, base, sets[i], pfe_quantile, threads, &outs[i], nullptr, 0.99, 10.0 / 250.0, nullptr, &rep,
             &sim);
    gTiming.total_s = secondsSince(t0);
  });
}

// Nested conditional initial margin of netting set `set_index` of `sets` (see nestedIm): per
// reporting date E[D IM], E[IM] and the 5 / 50 / 95% quantiles of IM over the paths. sample_path
// >= 0: also the inner samples of (sample_path, sample_rep) into samples (inner values) and that
// IM into sample_im.
int32_t ccr_oracle_nested_im(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                             const double* base_grid, int32_t base_size, const ccr_netting_set* sets, int32_t set_index,
                             int32_t inner_paths, double quantile, double mpor, int32_t threads, double* discounted,
                             double* mean, double* band05, double* band50, double* band95, int32_t capacity,
                             int32_t* num_dates, int64_t sample_path, int32_t sample_rep, double* samples,
                             double* sample_im, char* err, int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !sets || set_index < 0) invalid("null argument");
    if (inner_paths < 2 || inner_paths > 1024) invalid("nested initial margin: 2 to 1024 inner paths");
    if (static_cast<double>(inner_paths - 1) - std::floor(quantile * (inner_paths - 1)) >= 15.0)
      invalid("nested initial margin: the quantile needs more inner paths");
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    const NestedResult r = nestedIm(g, *config, base, sets[set_index], set_index, inner_paths, quantile, mpor, threads,
                                    sample_path, sample_rep);
    const int32_t n = static_cast<int32_t>(r.mean.size());
    if (num_dates) *num_dates = n;
    if (n > capacity) throw Failure(CCR_ERR_OUT_OF_RANGE, "capacity too small");
    for (int32_t i = 0; i < n; ++i) {
      if (discounted) discounted[i] = r.discounted[static_cast<std::size_t>(i)];
      if (mean) mean[i] = r.mean[static_cast<std::size_t>(i)];
      if (band05) band05[i] = r.b05[static_cast<std::size_t>(i)];
      if (band50) band50[i] = r.b50[static_cast<std::size_t>(i)];
      if (band95) band95[i] = r.b95[static_cast<std::size_t>(i)];
    }
    if (samples)
      for (std::size_t i = 0; i < r.samples.size(); ++i) samples[i] = r.samples[i];
    if (sample_im) *sample_im = r.sampleIm;
  });
}

// initialMarginProfile of netting set `index` of a book (its values on the book's simulation grid).
int32_t ccr_oracle_book_initial_margin(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                       const double* base_grid, int32_t base_size, const ccr_netting_set* sets,
                                       int32_t num_sets, int32_t index, double quantile_level, double mpor,
                                       int32_t threads, ccr_exposure_out* out, double* margin, int32_t capacity,
                                       int32_t* size, char* err, int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !sets || !out || !margin || index < 0 || index >= num_sets) invalid("bad argument");
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    std::vector<std::vector<Trade>> trades(static_cast<std::size_t>(num_sets));
    Vec r, v;
    for (int i = 0; i < num_sets; ++i) {
      for (int k = 0; k < sets[i].num_trades; ++k) trades[static_cast<std::size_t>(i)].push_back(makeTrade(sets[i].trades[k], g));
      const Grid gr = reportingGrid(base, trades[static_cast<std::size_t>(i)]);
      r.insert(r.end(), gr.t.begin(), gr.t.end());
    }
    const Grid rep(r);
    for (int i = 0; i < num_sets; ++i) {
      const Grid gs = simulationGrid(rep, trades[static_cast<std::size_t>(i)], sets[i]);
      v.insert(v.end(), gs.t.begin(), gs.t.end());
    }
    const Grid sim(v);
    Vec im;
    runSet(g, *config, base, sets[index], 0.95, threads, out, &im, quantile_level, mpor, nullptr, &rep, &sim);
    if (size) *size = static_cast<int32_t>(im.size());
    if (capacity < static_cast<int32_t>(im.size())) invalid("capacity smaller than the simulation grid");
    std::copy(im.begin(), im.end(), margin);
  });
}

// Per-trade statistics of one netting set next to its exposure profile: stats[(i * 5 + m) *
// capacity + r] for trade i, statistic m, reporting date r.
int32_t ccr_oracle_trade_statistics(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                    const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                    int32_t threads, ccr_exposure_out* out, double* stats, int32_t capacity, char* err,
                                    int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !set || !out || !stats) invalid("null argument");
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    std::vector<Vec> ts;
    runSet(g, *config, base, *set, 0.95, threads, out, nullptr, 0.99, 10.0 / 250.0, &ts);
    const std::size_t nR = static_cast<std::size_t>(out->num_dates);
    if (capacity < out->num_dates) invalid("capacity smaller than the reporting grid");
    for (std::size_t i = 0; i < ts.size(); ++i)
      for (int m = 0; m < kTradeStatCount; ++m)
        for (std::size_t r = 0; r < nR; ++r)
          stats[(i * kTradeStatCount + static_cast<std::size_t>(m)) * static_cast<std::size_t>(capacity) + r] =
              ts[i][static_cast<std::size_t>(m) * nR + r];
  });
}

// initialMarginProfile of one netting set (its simulation grid) next to its exposure profile.
int32_t ccr_oracle_initial_margin(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                                  const double* base_grid, int32_t base_size, const ccr_netting_set* set,
                                  double quantile_level, double mpor, int32_t threads, ccr_exposure_out* out,
                                  double* margin, int32_t capacity, int32_t* size, char* err, int32_t err_capacity) {
  return guard(err, err_capacity, [&] {
    if (!generator || !config || !set || !out || !margin) invalid("null argument");
    const Generator g(*generator);
    const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
    Vec im;
    runSet(g, *config, base, *set, 0.95, threads, out, &im, quantile_level, mpor);
    if (size) *size = static_cast<int32_t>(im.size());
    if (capacity < static_cast<int32_t>(im.size())) invalid("capacity smaller than the simulation grid");
    std::copy(im.begin(), im.end(), margin);
  });
}

int32_t ccr_oracle_last_timing(ccr_oracle_timing* out) {
  *out = gTiming;
  return CCR_OK;
}

int32_t ccr_oracle_time_grid(const double* times, int32_t n, double* out, int32_t capacity, int32_t* size) {
  return writeGrid(Grid(Vec(times, times + std::max(0, n))), out, capacity, size);
}

int32_t ccr_oracle_uniform_grid(double horizon, int32_t steps, double* out, int32_t capacity, int32_t* size) {
  return guard(nullptr, 0, [&] { writeGrid(Grid::uniform(horizon, static_cast<std::size_t>(std::max(0, steps))), out, capacity, size); });
}

int32_t ccr_oracle_lagged_grid(const double* grid, int32_t n, double lag, double* out, int32_t capacity, int32_t* size) {
  return writeGrid(Grid(Vec(grid, grid + std::max(0, n))).lagged(lag), out, capacity, size);
}

int32_t ccr_oracle_index_at_or_before(const double* grid, int32_t n, double t) {
  Grid g;
  g.t.assign(grid, grid + n);
  return static_cast<int32_t>(g.indexAtOrBefore(t));
}

int32_t ccr_oracle_reporting_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set, double* out,
                                  int32_t capacity, int32_t* size) {
  const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
  return writeGrid(reportingGrid(base, tradesOf(set, nullptr)), out, capacity, size);
}

int32_t ccr_oracle_simulation_grid(const double* base_grid, int32_t base_size, const ccr_netting_set* set, double* out,
                                   int32_t capacity, int32_t* size) {
  const Grid base(Vec(base_grid, base_grid + std::max(0, base_size)));
  return writeGrid(simulationGrid(base, tradesOf(set, nullptr), *set), out, capacity, size);
}

void ccr_oracle_normals(uint64_t seed, int64_t n, double* out) {
  std::mt19937_64 engine(seed);
  std::normal_distribution<double> normal(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i) out[i] = normal(engine);
}

void ccr_oracle_mt19937_64(uint64_t seed, int64_t n, uint64_t* out) {
  std::mt19937_64 engine(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = engine();
}

double ccr_oracle_discount(const ccr_yield_curve* curve, double t) { return Curve(*curve).discount(t); }
double ccr_oracle_instantaneous_forward(const ccr_yield_curve* curve, double t) {
  return Curve(*curve).instantaneousForward(t);
}
double ccr_oracle_survival(const ccr_credit_curve* curve, double t) { return Credit(*curve).survival(t); }
double ccr_oracle_zero_bond(const ccr_hull_white* hw, double t, double T, double x) { return HW(*hw).zeroBond(t, T, x); }
void ccr_oracle_hw_evolve(const ccr_hull_white* hw, double t0, double t1, double x0, double z1, double z2, double drift,
                          double* x1, double* integrated_rate) {
  HW(*hw).evolve(t0, t1, x0, z1, z2, drift, *x1, *integrated_rate);
}
double ccr_oracle_cir_survival(const ccr_cir_intensity* cir, double t) { return Cir(*cir).cirSurvival(t); }
double ccr_oracle_quantile(const double* values, int64_t n, double p) {
  double r = std::nan("");
  guard(nullptr, 0, [&] { r = quantile(Vec(values, values + n), p); });
  return r;
}
int32_t ccr_oracle_cholesky(const double* a, int32_t n, double* l) {
  return guard(nullptr, 0, [&] {
    Mat m(static_cast<std::size_t>(n), static_cast<std::size_t>(n));
    std::memcpy(m.d.data(), a, sizeof(double) * static_cast<std::size_t>(n * n));
    const Mat r = cholesky(m);
    std::memcpy(l, r.d.data(), sizeof(double) * static_cast<std::size_t>(n * n));
  });
}
int32_t ccr_oracle_solve(const double* a, const double* b, int32_t n, double* x) {
  return guard(nullptr, 0, [&] {
    Mat m(static_cast<std::size_t>(n), static_cast<std::size_t>(n));
    std::memcpy(m.d.data(), a, sizeof(double) * static_cast<std::size_t>(n * n));
    const Vec r = solve(m, Vec(b, b + n));
    std::copy(r.begin(), r.end(), x);
  });
}
int32_t ccr_oracle_swap_rate_annuity(const ccr_hull_white* hw, const double* sched, int32_t n, double t, double x,
                                     double* rate, double* annuity) {
  return guard(nullptr, 0, [&] {
    if (n < 2) invalid("swap rate: the schedule needs two dates");
    swapRateAnnuityAt(Vec(sched, sched + n), HW(*hw), t, x, *rate, *annuity);
  });
}

int32_t ccr_oracle_sa_regression(const double* u1, const double* u2, const double* y, const uint8_t* mask, int64_t n,
                                 int32_t degree, double* coefficients, double* params) {
  return guard(nullptr, 0, [&] {
    SaReg reg(static_cast<std::size_t>(std::max(0, degree)));
    std::vector<char> m;
    if (mask) m.assign(mask, mask + n);
    reg.fit(Vec(u1, u1 + n), Vec(u2, u2 + n), Vec(y, y + n), mask ? &m : nullptr);
    for (int i = 0; i < 10; ++i) coefficients[i] = i < static_cast<int>(reg.coef.size()) ? reg.coef[static_cast<std::size_t>(i)] : 0.0;
    params[0] = reg.m1;
    params[1] = reg.s1;
    params[2] = reg.m2;
    params[3] = reg.s2;
    params[4] = static_cast<double>(reg.deg);
  });
}

int32_t ccr_oracle_set_regression_basis(int32_t basis) {
  if (basis < -1 || basis > 1) return CCR_ERR_INVALID_ARGUMENT;
  gBasis = basis;
  return CCR_OK;
}

int32_t ccr_oracle_regression(const double* x, const double* y, const uint8_t* mask, int64_t n, int32_t degree,
                              double* coefficients, double* center, double* scale, int32_t* effective_degree) {
  return guard(nullptr, 0, [&] {
    PolyReg reg(static_cast<std::size_t>(degree));
    std::vector<char> m;
    if (mask) m.assign(mask, mask + n);
    reg.fit(Vec(x, x + n), Vec(y, y + n), mask ? &m : nullptr);
    for (int32_t d = 0; d <= degree; ++d) coefficients[d] = d < static_cast<int32_t>(reg.coef.size()) ? reg.coef[d] : 0.0;
    *center = reg.center;
    *scale = reg.scale;
    *effective_degree = static_cast<int32_t>(reg.coef.size()) - 1;
  });
}
int32_t ccr_oracle_collateral(const ccr_collateral* csa, const double* values, const double* grid, int32_t paths,
                              int32_t dates, const double* unit_price, double* out) {
  return guard(nullptr, 0, [&] {
    Mat v(static_cast<std::size_t>(paths), static_cast<std::size_t>(dates));
    std::copy(values, values + paths * dates, v.d.begin());
    Grid g;
    g.t.assign(grid, grid + dates);
    Mat pr;
    if (unit_price) {
      pr = Mat(static_cast<std::size_t>(paths), static_cast<std::size_t>(dates));
      std::copy(unit_price, unit_price + paths * dates, pr.d.begin());
    }
    const Mat c = collateral(*csa, v, g, unit_price ? &pr : nullptr);
    std::copy(c.d.begin(), c.d.end(), out);
  });
}
int32_t ccr_oracle_profile(const double* values, const double* deflator, int32_t paths, int32_t dates,
                           double pfe_quantile, double* expected_value, double* ee, double* ene, double* pfe,
                           double* dee, double* dene) {
  return guard(nullptr, 0, [&] {
    const double n = static_cast<double>(paths);
    Vec pos(static_cast<std::size_t>(paths));
    for (int32_t j = 0; j < dates; ++j) {
      double sum = 0.0, e = 0.0, ne = 0.0, de = 0.0, dne = 0.0;
      for (int32_t p = 0; p < paths; ++p) {
        const double v = values[p * dates + j], d = deflator[p * dates + j];
        sum += v;
        pos[static_cast<std::size_t>(p)] = std::max(v, 0.0);
        e += pos[static_cast<std::size_t>(p)];
        ne += std::min(v, 0.0);
        de += d * pos[static_cast<std::size_t>(p)];
        dne += d * std::min(v, 0.0);
      }
      expected_value[j] = sum / n;
      ee[j] = e / n;
      ene[j] = ne / n;
      dee[j] = de / n;
      dene[j] = dne / n;
      pfe[j] = quantile(pos, pfe_quantile);
    }
  });
}

// ExposureProfile::timeAverage: src/exposure/exposure_profile.cpp:51-60
static double timeAverage(const double* t, const double* v, int32_t n, double horizon) {
  double integral = 0.0, end = 0.0;
  for (int32_t j = 1; j < n && t[j] <= horizon + 1e-10; ++j) {
    integral += v[j] * (t[j] - t[j - 1]);
    end = t[j];
  }
  return end > 0.0 ? integral / end : (n == 0 ? 0.0 : v[0]);
}
double ccr_oracle_epe(const double* times, const double* ee, int32_t n, double horizon) {
  return timeAverage(times, ee, n, horizon);
}
double ccr_oracle_effective_epe(const double* times, const double* ee, int32_t n, double horizon) {
  Vec eff(static_cast<std::size_t>(n));
  double running = 0.0;
  for (int32_t j = 0; j < n; ++j) {
    running = std::max(running, ee[j]);
    eff[static_cast<std::size_t>(j)] = running;
  }
  return timeAverage(times, eff.data(), n, horizon);
}
double ccr_oracle_unilateral_cva(const double* times, const double* dee, int32_t n, const ccr_credit_curve* cpty,
                                 double recovery) {
  const Credit c(*cpty);
  Vec terms(static_cast<std::size_t>(n), 0.0);
  for (int32_t i = 1; i < n; ++i) {
    const double ee = 0.5 * (dee[i - 1] + dee[i]);
    terms[static_cast<std::size_t>(i)] = (1.0 - recovery) * ee * c.defaultProbability(times[i - 1], times[i]);
  }
  double cva = 0.0;
  for (double t : terms) cva += t;
  return cva;
}
void ccr_oracle_bilateral_cva(const double* t, const double* epe, const double* ene, int32_t n,
                              const ccr_credit_curve* cpty, double rc, const ccr_credit_curve* own, double ro,
                              double* cva, double* dva) {
  const Credit c(*cpty), o(*own);
  double a = 0.0, b = 0.0;
  for (int32_t i = 1; i < n; ++i) {
    const double mid = 0.5 * (t[i - 1] + t[i]);
    a += (1.0 - rc) * 0.5 * (epe[i - 1] + epe[i]) * o.survival(mid) * c.defaultProbability(t[i - 1], t[i]);
    b += (1.0 - ro) * 0.5 * (-ene[i - 1] - ene[i]) * c.survival(mid) * o.defaultProbability(t[i - 1], t[i]);
  }
  *cva = a;
  *dva = b;
}

// xva (src/xva/xva.cpp:16-96): CVA/DVA first-to-default with the other party's survival at the
// mid-points; FCA, FBA, MVA and KVA as survival-weighted trapezoid integrals (survivals of both
// parties at the mid-points); FBA on the magnitude of ENE*. out = cva, dva, fca, fba, mva, kva.
// params = borrow, lend, margin funding, hurdle, capital ratio, risk weight, alpha.
void ccr_oracle_xva(const double* t, const double* dee, const double* dene, int32_t n, const double* dim,
                    const ccr_credit_curve* cpty, double rc, const ccr_credit_curve* own, double ro,
                    const double* params, double* out) {
  const Credit c(*cpty), o(*own);
  auto weighted = [&](const Vec& f) {
    double total = 0.0;
    for (int32_t i = 1; i < n; ++i) {
      const double mid = 0.5 * (t[i - 1] + t[i]);
      total += 0.5 * (f[static_cast<std::size_t>(i - 1)] + f[static_cast<std::size_t>(i)]) * (t[i] - t[i - 1]) *
               c.survival(mid) * o.survival(mid);
    }
    return total;
  };
  Vec ee(dee, dee + n), mag(static_cast<std::size_t>(n));
  for (int32_t i = 0; i < n; ++i) mag[static_cast<std::size_t>(i)] = -dene[i];
  double cva = 0.0, dva = 0.0;
  for (int32_t i = 1; i < n; ++i) {
    const double mid = 0.5 * (t[i - 1] + t[i]);
    cva += (1.0 - rc) * 0.5 * (dee[i - 1] + dee[i]) * o.survival(mid) * c.defaultProbability(t[i - 1], t[i]);
    dva += (1.0 - ro) * 0.5 * (-dene[i - 1] - dene[i]) * c.survival(mid) * o.defaultProbability(t[i - 1], t[i]);
  }
  const double w = weighted(ee);
  out[0] = cva;
  out[1] = dva;
  out[2] = params[0] * w;
  out[3] = params[1] * weighted(mag);
  out[4] = dim ? params[2] * weighted(Vec(dim, dim + n)) : 0.0;
  out[5] = params[3] * params[4] * params[5] * params[6] * w;
}

int32_t ccr_oracle_bermudan(const ccr_scenario_generator* generator, const ccr_simulation_config* config,
                            const double* grid, int32_t grid_size, const ccr_trade* trade, double* price,
                            int32_t* exercise_index) {
  return guard(nullptr, 0, [&] {
    const Generator g(*generator);
    const Trade tr = makeTrade(*trade, g);
    const Grid gr(Vec(grid, grid + grid_size));
    const Scenarios s = g.generate(gr, static_cast<std::size_t>(config->num_paths), config->seed, config->antithetic != 0, nullptr);
    const BermudanValuation v = bermudan(tr, g, s);
    *price = v.price;
    if (exercise_index)
      for (std::size_t p = 0; p < s.P; ++p) exercise_index[p] = static_cast<int32_t>(v.exerciseIndex[p]);
  });
}

}  // extern "C"
