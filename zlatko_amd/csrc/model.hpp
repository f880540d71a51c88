// model.hpp -- host-side market, model and grid arithmetic of the exposure plan (FP64).
//
// These are the path-independent pieces of the reference models, evaluated once per date on
// the host so that the kernels only see per-date coefficient tables:
//   TimeGrid         src/core/time_grid.cpp:10-66  (dedupe at 1e-9, lagged dates, indexAtOrBefore)
//   YieldCurve       src/market/yield_curve.cpp:20-45
//   CreditCurve      src/market/credit_curve.cpp:23-41
//   HullWhite1F      src/models/hull_white.cpp:19-68 (B, V, zero-bond factor, exact step coefficients)
//   CirIntensity     src/models/cir_intensity.cpp:25-47 (analytic survival, CIR++ shift)
// Grid and index arithmetic must be bit-identical to the reference (it decides which dates exist
// and which date a fixing or a margin call reads), so it follows the same comparisons.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ccr_hip.h"

namespace ccrhip {

struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int32_t code, const std::string& m) { throw Error(code, m); }
[[noreturn]] inline void invalid(const std::string& m) { fail(CCR_ERR_INVALID_ARGUMENT, m); }

constexpr double kGridTolerance = 1e-9;  // TimeGrid de-duplication and look-up tolerance
constexpr double kPaidTolerance = 1e-10; // swap periods paid / fixed, forward expiry

struct TimeGrid {
  std::vector<double> t{0.0};
  TimeGrid() = default;
  explicit TimeGrid(std::vector<double> v) {
    v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return x < 0.0 || !std::isfinite(x); }), v.end());
    v.push_back(0.0);
    std::sort(v.begin(), v.end());
    t.clear();
    for (double x : v)
      if (t.empty() || x - t.back() > kGridTolerance) t.push_back(x);
  }
  TimeGrid merged(const std::vector<double>& extra) const {
    std::vector<double> v = t;
    v.insert(v.end(), extra.begin(), extra.end());
    return TimeGrid(std::move(v));
  }
  TimeGrid lagged(double lag) const {
    if (lag <= 0.0) return *this;
    std::vector<double> v = t;
    for (double x : t)
      if (x > lag) v.push_back(x - lag);
    return TimeGrid(std::move(v));
  }
  std::size_t size() const { return t.size(); }
  double operator[](std::size_t i) const { return t[i]; }
  std::size_t indexAtOrBefore(double x) const {
    const auto it = std::upper_bound(t.begin(), t.end(), x + kGridTolerance);
    return it == t.begin() ? 0 : static_cast<std::size_t>(it - t.begin()) - 1;
  }
  std::optional<std::size_t> find(double x, double tol = kGridTolerance) const {
    const std::size_t i = indexAtOrBefore(x);
    if (std::fabs(t[i] - x) <= tol) return i;
    if (i + 1 < t.size() && std::fabs(t[i + 1] - x) <= tol) return i + 1;
    return std::nullopt;
  }
};

struct Curve {
  std::vector<double> t, r;
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
};

struct Credit {
  std::vector<double> p, h;
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
};

// G1++ pricing factors: P(t, T | x) = A(t, T) exp(-B(t, T) x) with
// A = P(0,T)/P(0,t) exp(1/2 (V(t,T) - V(0,T) + V(0,t))).
struct HullWhite {
  Curve curve;
  double a = 0.0, s = 0.0;
  HullWhite() = default;
  explicit HullWhite(const ccr_hull_white& m) : curve(m.curve), a(m.mean_reversion), s(m.volatility) {
    if (a <= 1e-8) invalid("HullWhite1F: mean reversion must be positive");
    if (s < 0.0) invalid("HullWhite1F: volatility must be non-negative");
  }
  double B(double t, double T) const { return (1.0 - std::exp(-a * (T - t))) / a; }
  double V(double t, double T) const {
    const double tau = T - t;
    const double e1 = std::exp(-a * tau);
    return s * s / (a * a) * (tau + 2.0 / a * e1 - 0.5 / a * e1 * e1 - 1.5 / a);
  }
  double A(double t, double T) const {
    return curve.discount(T) / curve.discount(t) * std::exp(0.5 * (V(t, T) - V(0.0, T) + V(0.0, t)));
  }
  // Exact step coefficients of (x, int x) over [t0, t1] and the integral of phi.
  struct Step {
    double e, sdX, c1, c2, xc, iphi;
  };
  Step step(double t0, double t1) const {
    const double dt = t1 - t0;
    Step st{};
    st.e = std::exp(-a * dt);
    const double s2 = s * s, e = st.e;
    const double varX = s2 * (1.0 - e * e) / (2.0 * a);
    const double varI = s2 / (a * a) * (dt - 2.0 * (1.0 - e) / a + (1.0 - e * e) / (2.0 * a));
    const double cov = s2 / (2.0 * a * a) * (1.0 - e) * (1.0 - e);
    st.sdX = std::sqrt(varX);
    st.c1 = 0.0;
    st.c2 = std::sqrt(std::max(varI, 0.0));
    if (st.sdX > 0.0) {
      st.c1 = cov / st.sdX;
      st.c2 = std::sqrt(std::max(varI - st.c1 * st.c1, 0.0));
    }
    st.xc = (1.0 - e) / a;
    const double g2 = dt - 2.0 / a * (std::exp(-a * t0) - std::exp(-a * t1)) +
                      0.5 / a * (std::exp(-2.0 * a * t0) - std::exp(-2.0 * a * t1));
    st.iphi = std::log(curve.discount(t0) / curve.discount(t1)) + s2 / (2.0 * a * a) * g2;
    return st;
  }
};

struct CirModel {
  Credit market;
  double kappa = 0, theta = 0, xi = 0, y0 = 0;
  int sub = 1;
  CirModel() = default;
  explicit CirModel(const ccr_cir_intensity& m)
      : market(m.market), kappa(m.kappa), theta(m.theta), xi(m.xi), y0(m.y0), sub(m.substeps) {
    if (kappa <= 0.0 || theta < 0.0 || xi < 0.0 || y0 < 0.0) invalid("CirIntensity: invalid parameters");
    if (sub <= 0) invalid("CirIntensity: substeps must be positive");
  }
  double cirSurvival(double t) const {
    if (t <= 0.0) return 1.0;
    if (xi == 0.0) return std::exp(-(theta * t + (y0 - theta) * (1.0 - std::exp(-kappa * t)) / kappa));
    const double h = std::sqrt(kappa * kappa + 2.0 * xi * xi);
    const double eh = std::exp(h * t) - 1.0;
    const double den = 2.0 * h + (kappa + h) * eh;
    const double logA = (2.0 * kappa * theta / (xi * xi)) * std::log(2.0 * h * std::exp(0.5 * (kappa + h) * t) / den);
    return std::exp(logA - 2.0 * eh / den * y0);
  }
  // ln(S_CIR(t) / S_market(t)): the CIR++ shift integral.
  double shift(double t) const { return std::log(cirSurvival(t) / market.survival(t)); }
};

// Lower-triangular Cholesky factor (src/core/matrix.cpp:49-68 semantics).
inline std::vector<double> cholesky(const std::vector<double>& a, std::size_t n) {
  std::vector<double> l(n * n, 0.0);
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = 0; j <= i; ++j) {
      if (std::fabs(a[i * n + j] - a[j * n + i]) > 1e-10) invalid("cholesky: matrix must be symmetric");
      double sum = a[i * n + j];
      for (std::size_t k = 0; k < j; ++k) sum -= l[i * n + k] * l[j * n + k];
      if (i == j) {
        if (sum <= 0.0) invalid("cholesky: matrix is not positive definite");
        l[i * n + i] = std::sqrt(sum);
      } else {
        l[i * n + j] = sum / l[j * n + j];
      }
    }
  return l;
}

}  // namespace ccrhip
