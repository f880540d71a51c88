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
// The same from the state x: z = (x - center) / scale.
__host__ __device__ inline double regPredict(double x, double center, double scale, int deg, const double* coef) {
  return regPredictZ((x - center) / scale, deg, coef);
}

}  // namespace ccrhip
