// amc.hpp -- host/device arithmetic of the American Monte Carlo regressions.
//
// PolynomialRegression (src/core/regression.cpp:29-90): standardise the regressor (mean,
// population sd), drop to degree 0 for a degenerate regressor, lower the degree for small samples,
// fit on the monomials 1, z, z^2, z^3 by normal equations with a ridge of 1e-12 n, Horner predict.
// solveLinearSystem (src/core/matrix.cpp:70-96): Gaussian elimination with partial pivoting.
// The moments themselves are reduced by the GPU kernels (fixed order, FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "launch.hpp"

namespace ccrhip {

struct RegParams {
  double n = 0.0, center = 0.0, scale = 1.0;
  int deg = -1;  // -1: no sample (prediction 0); 0: sample mean; >= 1: polynomial degree
};

// Standardisation and effective degree from n, sum x, sum x^2 (regression.cpp:41-62).
__host__ __device__ inline RegParams regParams(double n, double sx, double sxx, int degree) {
  RegParams r;
  r.n = n;
  if (n == 0.0) return r;
  r.center = sx / n;
  const double var = fmax(sxx / n - r.center * r.center, 0.0);
  r.scale = sqrt(var);
  int deg = degree;
  if (r.scale < 1e-12 * fmax(1.0, fabs(r.center))) deg = 0;
  const long long nn = static_cast<long long>(n);
  if (nn < 2LL * (deg + 1)) deg = nn >= 4 ? static_cast<int>(nn / 2) - 1 : 0;
  if (deg == 0) r.scale = 1.0;
  r.deg = deg;
  return r;
}

// Design row 1, z, z^2, z^3 as the reference builds it (repeated multiplication).
__host__ __device__ inline void regRow(double z, double* phi) {
  phi[0] = 1.0;
  phi[1] = z;
  phi[2] = z * z;
  phi[3] = phi[2] * z;
}

// Index of (a, b), b <= a, in a packed lower triangle.
__host__ __device__ inline int triIndex(int a, int b) { return a * (a + 1) / 2 + b; }

// Coefficients from the packed Gram (lower triangle of X'X), X'y and sum y: normal equations
// with the ridge, solved by elimination with partial pivoting. Returns false when singular.
__host__ __device__ inline bool regSolve(const RegParams& p, const double* gram, const double* xty, double sy,
                                         double* coef) {
  coef[0] = coef[1] = coef[2] = coef[3] = 0.0;
  if (p.deg < 0) return true;
  if (p.deg == 0) {
    coef[0] = sy / p.n;
    return true;
  }
  const int k = p.deg + 1;
  double a[4][4], b[4];
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j <= i; ++j) a[i][j] = a[j][i] = gram[triIndex(i, j)];
    a[i][i] += 1e-12 * p.n;
    b[i] = xty[i];
  }
  for (int col = 0; col < k; ++col) {
    int piv = col;
    for (int r = col + 1; r < k; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    if (fabs(a[piv][col]) < 1e-300) return false;
    if (piv != col) {
      for (int c = 0; c < k; ++c) {
        const double t = a[col][c];
        a[col][c] = a[piv][c];
        a[piv][c] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < k; ++r) {
      const double f = a[r][col] / a[col][col];
      if (f == 0.0) continue;
      for (int c = col; c < k; ++c) a[r][c] -= f * a[col][c];
      b[r] -= f * b[col];
    }
  }
  for (int i = k - 1; i >= 0; --i) {
    double s = b[i];
    for (int c = i + 1; c < k; ++c) s -= a[i][c] * coef[c];
    coef[i] = s / a[i][i];
  }
  return true;
}

// PolynomialRegression::predict (regression.cpp:85-90), from the standardised state z.
__host__ __device__ inline double regPredictZ(double z, int deg, const double* coef) {
  double r = 0.0;
  for (int d = deg < 0 ? 0 : deg; d >= 0; --d) r = r * z + coef[d];
  return r;
}
// The same for coefficients stored as a cubic with zeros above the degree (the valuation's
// continuation records): the zero leading steps of Horner's rule leave the next coefficient
// exactly (0 z + c = c), so the value is that of the degree's own recurrence, with no loop over
// a per-record degree.
__host__ __device__ inline double regPredictCubic(double z, const double* coef) {
  return ((coef[3] * z + coef[2]) * z + coef[1]) * z + coef[0];
}
// The same from the state x: z = (x - center) / scale.
__host__ __device__ inline double regPredict(double x, double center, double scale, int deg, const double* coef) {
  return regPredictZ((x - center) / scale, deg, coef);
}

// ---------------------------------------------------------------- regression engine (basis choice)
// Terms of a total-degree-d basis in two variables.
__host__ __device__ inline int basisTerms2(int deg) { return (deg + 1) * (deg + 2) / 2; }
// Index of (i, j), i <= j < 11, in the upper triangle of the Gram of [phi | y] (y is column 10).
__host__ __device__ inline int gramIndex(int i, int j) { return i * (21 - i) / 2 + j; }

// Standardisation and effective degree of a record from its moments m[kBasisMom] and the
// swaption's degree (the total degree, at most 3, of the swap-rate / annuity basis): writes n, m1, sd1, m2, sd2, degree into rp[0..5] and, for
// degree 0, the sample mean of y into rp[6]. Degree -1: no sample (prediction 0).
__host__ __device__ inline void basisParams(int basis, int degree, const double* m, double* rp) {
  for (int i = 0; i < kBasisStride; ++i) rp[i] = 0.0;
  const double n = m[0];
  rp[0] = n;
  rp[2] = rp[4] = 1.0;
  if (n == 0.0) {
    rp[5] = -1.0;
    return;
  }
  int deg;
  if (basis == kBasisState) {
    const RegParams r = regParams(n, m[1], m[2], degree);
    rp[1] = r.center;
    rp[2] = r.scale;
    deg = r.deg;
  } else {
    // The state basis's rules on each variable: population deviations; a degenerate variable
    // drops the fit to the mean; the total degree falls until n >= 2 x terms. A constant
    // regressor's one-pass variance is rounding of ~1e-16 E[u^2], so a deviation below 1e-6 of
    // the root mean square counts as degenerate.
    const double m1 = m[1] / n, m2 = m[3] / n;
    const double sd1 = sqrt(fmax(m[2] / n - m1 * m1, 0.0)), sd2 = sqrt(fmax(m[4] / n - m2 * m2, 0.0));
    deg = degree < 0 ? 0 : (degree > 3 ? 3 : degree);
    if (!(sd1 > kSaDegenerate * sqrt(m[2] / n)) || !(sd2 > kSaDegenerate * sqrt(m[4] / n))) deg = 0;
    while (deg > 0 && n < 2.0 * basisTerms2(deg)) --deg;
    rp[1] = m1;
    rp[3] = m2;
    rp[2] = deg > 0 ? sd1 : 1.0;
    rp[4] = deg > 0 ? sd2 : 1.0;
  }
  rp[5] = static_cast<double>(deg);
  if (deg == 0) rp[6] = m[5] / n;
}

// Number of terms of a record's fit.
__host__ __device__ inline int basisCount(int basis, int deg) {
  if (deg < 0) return 0;
  return basis == kBasisState ? deg + 1 : basisTerms2(deg);
}

// Design row of a sample with raw features (u1, u2): the state basis's powers of z (repeated
// multiplication, as the reference builds them) or the monomials of (s, a).
__host__ __device__ inline void basisRow(int basis, const double* rp, double u1, double u2, double* phi) {
  if (basis == kBasisState) {
    regRow((u1 - rp[1]) / rp[2], phi);
    for (int i = 4; i < kBasisMax; ++i) phi[i] = 0.0;
    return;
  }
  const double s = (u1 - rp[1]) / rp[2], a = (u2 - rp[3]) / rp[4];
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

// Prediction of a fitted record: the state basis by Horner in z (PolynomialRegression::predict),
// the swap-rate / annuity basis as the sum of coefficient x term in term order.
__host__ __device__ inline double basisPredict(int basis, const double* rp, double u1, double u2) {
  const int deg = static_cast<int>(rp[5]);
  if (deg < 0) return 0.0;
  if (basis == kBasisState) return regPredictZ((u1 - rp[1]) / rp[2], deg, rp + 6);
  double phi[kBasisMax];
  basisRow(basis, rp, u1, u2, phi);
  const int k = basisCount(basis, deg);
  double r = 0.0;
  for (int i = 0; i < k; ++i) r += rp[6 + i] * phi[i];
  return r;
}

// Coefficients of a record from its Gram sums g[kBasisGram] (rp from basisParams). The state
// basis: normal equations with the ridge 1e-12 n, elimination with partial pivoting
// (matrix.cpp:70-96), as the reference. The swap-rate / annuity basis: in a one-factor model both
// regressors are functions of the one state, so its monomials are numerically dependent (Gram
// condition ~1e16); a ridge of kSaRidge x the Gram diagonal (besides 1e-12 n) makes the fit
// reproducible across summation orders and Gram precisions, and the system is solved by FP64
// Cholesky. Returns false when the factorisation fails (the fit falls back to the sample mean).
__host__ __device__ inline bool basisSolve(int basis, const double* g, double* rp) {
  const int deg = static_cast<int>(rp[5]);
  if (deg <= 0) return true;
  const int k = basisCount(basis, deg);
  double a[kBasisMax][kBasisMax], b[kBasisMax], x[kBasisMax];
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < k; ++j) a[i][j] = g[gramIndex(i < j ? i : j, i < j ? j : i)];
    a[i][i] += basis == kBasisState ? 1e-12 * rp[0] : 1e-12 * rp[0] + kSaRidge * g[gramIndex(i, i)];
    b[i] = g[gramIndex(i, 10)];
  }
  if (basis != kBasisState) {
    // a = L L^T (L in the lower triangle of a), then L z = b, L^T x = z.
    for (int j = 0; j < k; ++j) {
      double d = a[j][j];
      for (int m = 0; m < j; ++m) d -= a[j][m] * a[j][m];
      if (!(d > 0.0)) {
        for (int i = 6; i < kBasisStride; ++i) rp[i] = 0.0;
        rp[5] = 0.0;
        rp[6] = g[gramIndex(0, 10)] / rp[0];
        return false;
      }
      const double l = sqrt(d);
      a[j][j] = l;
      for (int i = j + 1; i < k; ++i) {
        double s = a[i][j];
        for (int m = 0; m < j; ++m) s -= a[i][m] * a[j][m];
        a[i][j] = s / l;
      }
    }
    for (int i = 0; i < k; ++i) {
      double s = b[i];
      for (int m = 0; m < i; ++m) s -= a[i][m] * x[m];
      x[i] = s / a[i][i];
    }
    for (int i = k - 1; i >= 0; --i) {
      double s = x[i];
      for (int m = i + 1; m < k; ++m) s -= a[m][i] * x[m];
      x[i] = s / a[i][i];
    }
    for (int i = 0; i < k; ++i) rp[6 + i] = x[i];
    return true;
  }
  for (int col = 0; col < k; ++col) {
    int piv = col;
    for (int r = col + 1; r < k; ++r)
      if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
    if (fabs(a[piv][col]) < 1e-300) return false;
    if (piv != col) {
      for (int c = 0; c < k; ++c) {
        const double t = a[col][c];
        a[col][c] = a[piv][c];
        a[piv][c] = t;
      }
      const double t = b[col];
      b[col] = b[piv];
      b[piv] = t;
    }
    for (int r = col + 1; r < k; ++r) {
      const double f = a[r][col] / a[col][col];
      if (f == 0.0) continue;
      for (int c = col; c < k; ++c) a[r][c] -= f * a[col][c];
      b[r] -= f * b[col];
    }
  }
  for (int i = k - 1; i >= 0; --i) {
    double s = b[i];
    for (int c = i + 1; c < k; ++c) s -= a[i][c] * x[c];
    x[i] = s / a[i][i];
  }
  for (int i = 0; i < k; ++i) rp[6 + i] = x[i];
  return true;
}

}  // namespace ccrhip
