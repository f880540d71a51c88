// path_body.hpp -- the per-path body of the fused exposure kernel, one Monte Carlo path.
//
// A __host__ __device__ template: ccr_path_kernel runs it with one path per lane on gfx950;
// the test-only CPU emulator runs the same code path by path to pin the plan tables against the
// FP64 oracle without a GPU. For one path, over the simulation dates:
//   replayed normals (sign-flipped for the antithetic mirror, scenario_generator.cpp:94-100)
//   -> w = L z (scenario_generator.cpp:118-124)
//   -> exact Hull-White steps of every rate (hull_white.cpp:40-68; foreign rates with the quanto
//      drift), deflator D *= exp(-int r) (scenario_generator.cpp:126-128)
//   -> lognormal assets (lognormal_asset.cpp:17-22), carry from a curve or a foreign rate
//   -> CIR++ with Brownian-bridge sub-steps and pathwise survival (cir_intensity.cpp:45-66)
//   -> on valued dates V = sum over rate groups of FX x (bond polynomial or merged bond terms
//      + fixed coupons) + deterministic-carry forwards
//   -> CSA (collateral.cpp:10-47) with the margin-call value read from a per-path ring
//   -> on reporting dates, the 9 statistics of ExposureProfile::compute and pathwiseCva.
//
// The model shape (rates NR, assets NA, credits NC) is a template parameter: with kExact the
// counts are exactly those, so register arrays and loops are sized at compile time; the generic
// instantiation takes the maxima and reads the counts from the plan.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "amc.hpp"
#include "launch.hpp"

namespace ccrhip {

// Fused multiply-adds whose named operand is wavefront-uniform (a constant, or a value loaded
// through the scalar cache): one v_fma_f64 that reads it from SGPRs. Left to itself the compiler
// selects the two-address v_fmac_f64 and first copies such an operand into the destination
// VGPRs (two v_mov_b32 per use) or keeps constants resident in VGPRs. Same rounding as fma().
// Only for operands that are uniform across the wavefront.
__host__ __device__ inline double fmaU3(double a, double b, double c) {  // c uniform
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}
__host__ __device__ inline double fmaU2(double a, double b, double c) {  // b uniform
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

// exp(a) for |a| < 700 in FP64 (no overflow / underflow handling: bond-term exponents are
// O(1)). Cody-Waite reduction a = n ln2 + r, |r| <= ln2/2, with the two-part ln2 of fdlibm's
// e_exp.c; exp(r) by its degree-13 Taylor polynomial (truncation < 5e-18 relative on the
// interval) in Horner form; 2^n by ldexp. Accurate to about 1 ulp.
__host__ __device__ inline double expBond(double a) {
  const double n = rint(a * 1.4426950408889634074);
  double r = fmaU2(n, -6.93147180369123816490e-01, a);
  r = fmaU2(n, -1.90821492927058770002e-10, r);
  double p = fmaU3(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);  // 1/13!, 1/12!
  p = fmaU3(p, r, 1.0 / 39916800.0);     // 1/11!
  p = fmaU3(p, r, 1.0 / 3628800.0);      // 1/10!
  p = fmaU3(p, r, 1.0 / 362880.0);       // 1/9!
  p = fmaU3(p, r, 1.0 / 40320.0);        // 1/8!
  p = fmaU3(p, r, 1.0 / 5040.0);         // 1/7!
  p = fmaU3(p, r, 1.0 / 720.0);          // 1/6!
  p = fmaU3(p, r, 1.0 / 120.0);          // 1/5!
  p = fmaU3(p, r, 1.0 / 24.0);           // 1/4!
  p = fmaU3(p, r, 1.0 / 6.0);            // 1/3!
  p = fmaU3(p, r, 0.5);
  p = fmaU3(p, r, 1.0);
  p = fmaU3(p, r, 1.0);
  return ldexp(p, static_cast<int>(n));
}

template <typename Real>
__host__ __device__ inline Real expr(Real v) {
  if constexpr (std::is_same<Real, double>::value) return exp(v);
  else return __expf(v);
}
// Exponential of a merged bond-term exponent.
template <typename Real>
__host__ __device__ inline Real expTerm(Real v) {
  if constexpr (std::is_same<Real, double>::value) return expBond(v);
  else return __expf(v);
}
// One Horner step h x + c of a merged-bond power series in the plan's precision (c uniform): the
// FP64 form reads c from SGPRs (fmaU3); FP32 plans hold FP32 coefficients.
template <typename Real>
__host__ __device__ inline Real hornerStep(Real h, Real x, Real c) {
  if constexpr (std::is_same<Real, double>::value) return fmaU3(h, x, c);
  else return fmaf(h, x, c);
}

// True when the predicate holds on every lane of the wavefront (the path itself on the host).
__host__ __device__ inline bool allLanes(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __all(p) != 0;
#else
  return p;
#endif
}
template <typename Real>
__host__ __device__ inline Real sqrtr(Real v) {
  if constexpr (std::is_same<Real, double>::value) return sqrt(v);
  else return __fsqrt_rn(v);
}

// normalCdf and blackFormula (include/ccr/core/math.hpp:15-31): N(x) = erfc(-x / sqrt 2) / 2;
// discount x sign x (F N(sign d1) - K N(sign d2)), the discounted intrinsic value when the
// total standard deviation, the forward or the strike degenerates. sign: +1 call, -1 put.
template <typename Real>
__host__ __device__ inline Real normalCdf(Real x) {
  if constexpr (std::is_same<Real, double>::value) return 0.5 * erfc(-x / 1.4142135623730951);
  else return 0.5f * erfcf(-x / 1.41421356f);
}
template <typename Real>
__host__ __device__ inline Real blackFormula(Real sign, Real forward, Real strike, Real stdDev, Real discount) {
  if (stdDev <= Real(1e-14) || forward <= Real(0) || strike <= Real(0)) {
    const Real intrinsic = sign * (forward - strike);
    return discount * (intrinsic > Real(0) ? intrinsic : Real(0));
  }
  Real lr;
  if constexpr (std::is_same<Real, double>::value) lr = log(forward / strike);
  else lr = logf(forward / strike);
  const Real d1 = lr / stdDev + Real(0.5) * stdDev;
  const Real d2 = d1 - stdDev;
  return discount * sign * (forward * normalCdf(sign * d1) - strike * normalCdf(sign * d2));
}

// Value of a TermRange (an instrument's terms on one date, indices relative to its set's tables)
// at domestic state x: the bond polynomial when every lane is within its bound, else the
// exponential terms; plus the fixed coupons, whose fixing states fixing(k) come from the states
// row k named in each coupon's fourth element.
template <typename Real, typename Fixing>
__host__ __device__ inline Real evalRange(const CCR_CONST TermRange& tr, Real x, cptr<Real> bonds, cptr<Real> fixed,
                                          cptr<Real> poly, Fixing&& fixing) {
  Real acc;
  if (tr.poly_len > 0 && allLanes(static_cast<double>(x < Real(0) ? -x : x) <= tr.xmax)) {
    // Horner over kPolyBatch coefficients per batch of scalar loads (the range is uniform
    // across the wavefront wherever this runs: one exercise or one valued date per workgroup).
    cptr<Real> pc = poly + tr.poly_beg;
    Real h = 0;
    for (int k = 0; k < tr.poly_len; k += kPolyBatch) {
      Real c[kPolyBatch];
#pragma unroll
      for (int u = 0; u < kPolyBatch; ++u) c[u] = pc[k + u];
#pragma unroll
      for (int u = 0; u < kPolyBatch; ++u) h = hornerStep<Real>(h, x, c[u]);
    }
    acc = h;
  } else {
    acc = static_cast<Real>(tr.c0);
    for (int i = tr.bond_beg; i < tr.bond_end; ++i) acc += bonds[2 * i] * expTerm(-bonds[2 * i + 1] * x);
  }
  for (int i = tr.fix_beg; i < tr.fix_end; ++i) {
    cptr<Real> f = fixed + 4 * static_cast<int64_t>(i);
    const Real xk = fixing(static_cast<int>(f[3]));
    acc += f[0] * expTerm(-f[1] * x - f[2] * xk);
  }
  return acc;
}

// Sum over the Bermudan swaptions of set S of position x value on date j of path p
// (bermudan_swaption.cpp:104-140): the underlying swap once exercised before the next exercise
// date (physical; 0 for cash), else max(continuation, 0). states holds x then D ([date][path]
// from S.state_off, D after state_len); bonds / fixed / poly are the set's tables. Only the
// swaptions live[lbeg, lend) are visited: the plan lists, per continuation date, those that can
// be nonzero on some path (the others add an exact 0).
template <typename Real, bool kSa = false>
__host__ __device__ inline Real bermudanSum(const CCR_CONST SetDev& S, int j, int64_t p, int64_t paths, cptr<double> cont,
                                            const int16_t* exidx, const double* states, cptr<Real> bonds,
                                            cptr<Real> fixed, cptr<Real> poly, cptr<LiveRec> live, int lbeg,
                                            int lend, const BasisView& eng) {
  const double* xrow = states + S.state_off;
  const double xd = xrow[static_cast<int64_t>(j) * paths + p];
  const Real x = static_cast<Real>(xd);
  auto fixing = [&](int k) { return static_cast<Real>(xrow[static_cast<int64_t>(k) * paths + p]); };
  Real V = 0;
  // The continuation fits of one date standardise x alike (the date's mean and deviation over
  // all paths): z is computed once and again only if a record holds other constants.
  double zCenter = 0.0, zScale = 0.0, z = 0.0;
  bool haveZ = false;
  for (int l = lbeg; l < lend; ++l) {
    // The swaption, its date record and that date's BermDate: the first 32 bytes of the live
    // record in two 16-byte scalar loads (member-wise reads become a chain of small loads).
#if defined(__HIP_DEVICE_COMPILE__)
    LiveRec lr;
    {
      cptr<int4> raw = reinterpret_cast<cptr<int4>>(live + l);
      const int4 h0 = raw[0], h1 = raw[1];
      lr.b = h0.x;
      lr.rec = h0.y;
      lr.ne = h0.z;
      lr.next_k = h0.w;
      lr.has_swap = h1.x;
      lr.position = __hiloint2double(h1.w, h1.z);
    }
#else
    const LiveRec lr = live[l];
#endif
    const bool exercised = exidx[static_cast<int64_t>(lr.b) * paths + p] < lr.next_k;
    Real v = 0;
    if constexpr (kSa) {
      // The continuation fit of this (swaption, date) on the swap rate and annuity of the tail
      // entered at the next exercise.
      const int rid = lr.next_k < lr.ne ? eng.rcont[lr.rec] : -1;
      if (rid >= 0) {
        const CCR_CONST RegRec& rr = eng.rrecs[rid];
        auto noFix = [](int) { return Real(0); };
        const double A = static_cast<double>(evalRange<Real>(eng.saterms[rr.ann], x, bonds, fixed, poly, noFix));
        const double F = static_cast<double>(evalRange<Real>(eng.saterms[rr.flt], x, bonds, fixed, poly, noFix));
        double rp[kBasisStride];
        for (int i = 0; i < kBasisStride; ++i) rp[i] = eng.rfit[static_cast<int64_t>(rid) * kBasisStride + i];
        const double pred = basisPredict(kBasisSwapRateAnnuity, rp, F / A, A);
        v = static_cast<Real>(pred > 0.0 ? pred : 0.0);
      }
    } else if (lr.next_k < lr.ne) {
      cptr<double> c = cont + static_cast<int64_t>(lr.rec) * kContStride;
      const double coef[4] = {c[2], c[3], c[4], c[5]};
      if (!haveZ || c[0] != zCenter || c[1] != zScale) {
        zCenter = c[0];
        zScale = c[1];
        z = (xd - zCenter) / zScale;
        haveZ = true;
      }
      const double pred = regPredictCubic(z, coef);  // the record's coefficients above its degree are 0
      v = static_cast<Real>(pred > 0.0 ? pred : 0.0);
    }
    if (lr.has_swap && !allLanes(!exercised)) {
      const Real sv = evalRange<Real>(live[l].swap, x, bonds, fixed, poly, fixing);
      if (exercised) v = sv;
    } else if (exercised) {
      v = 0;
    }
    V += static_cast<Real>(lr.position) * v;
  }
  return V;
}

// bermudanSum (state basis) for NP paths per thread, p[0..NP): the live records, the exercise
// decisions' scalar work and the date's standardisation are shared by the thread's paths, which
// divides the scalar instructions per path of the valuation kernel by NP and gives each record
// NP independent exercise-index loads and predictions in flight. Each path's value is the one
// bermudanSum computes; the uniform choices (evaluate the underlying, power series or
// exponential terms) are taken over the wavefront's lanes for each of its NP path slots.
template <typename Real, int NP>
__host__ __device__ inline void bermudanSumN(const CCR_CONST SetDev& S, int j, const int64_t (&p)[NP], int64_t paths,
                                             cptr<double> cont, const int16_t* exidx, const double* states,
                                             cptr<Real> bonds, cptr<Real> fixed, cptr<Real> poly, cptr<LiveRec> live,
                                             int lbeg, int lend, Real (&V)[NP]) {
  const double* xrow = states + S.state_off;
  double xd[NP], z[NP];
  Real x[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    xd[i] = xrow[static_cast<int64_t>(j) * paths + p[i]];
    x[i] = static_cast<Real>(xd[i]);
    z[i] = 0.0;
    V[i] = 0;
  }
  double zCenter = 0.0, zScale = 0.0;
  bool haveZ = false;
  for (int l = lbeg; l < lend; ++l) {
#if defined(__HIP_DEVICE_COMPILE__)
    LiveRec lr;
    {
      cptr<int4> raw = reinterpret_cast<cptr<int4>>(live + l);
      const int4 h0 = raw[0], h1 = raw[1];
      lr.b = h0.x;
      lr.rec = h0.y;
      lr.ne = h0.z;
      lr.next_k = h0.w;
      lr.has_swap = h1.x;
      lr.position = __hiloint2double(h1.w, h1.z);
    }
#else
    const LiveRec lr = live[l];
#endif
    bool exercised[NP], any = false;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      exercised[i] = exidx[static_cast<int64_t>(lr.b) * paths + p[i]] < lr.next_k;
      any = any || exercised[i];
    }
    Real v[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) v[i] = 0;
    if (lr.next_k < lr.ne) {
      cptr<double> c = cont + static_cast<int64_t>(lr.rec) * kContStride;
      const double coef[4] = {c[2], c[3], c[4], c[5]};
      if (!haveZ || c[0] != zCenter || c[1] != zScale) {
        zCenter = c[0];
        zScale = c[1];
#pragma unroll
        for (int i = 0; i < NP; ++i) z[i] = (xd[i] - zCenter) / zScale;
        haveZ = true;
      }
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const double pred = regPredictCubic(z[i], coef);
        v[i] = static_cast<Real>(pred > 0.0 ? pred : 0.0);
      }
    }
    if (lr.has_swap && !allLanes(!any)) {
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        const int64_t pi = p[i];
        auto fixing = [&](int k) { return static_cast<Real>(xrow[static_cast<int64_t>(k) * paths + pi]); };
        const Real sv = evalRange<Real>(live[l].swap, x[i], bonds, fixed, poly, fixing);
        if (exercised[i]) v[i] = sv;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NP; ++i)
        if (exercised[i]) v[i] = 0;
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) V[i] += static_cast<Real>(lr.position) * v[i];
  }
}

// Model shapes the path kernel is specialised for; everything else runs the generic shape.
template <int NR_, int NA_, int NC_, bool Exact_>
struct Shape {
  static constexpr int NR = NR_, NA = NA_, NC = NC_;
  static constexpr bool kExact = Exact_;
  static constexpr int kCorr = NR + NA + NC;  // correlated drivers
  static constexpr int kPre = kCorr + NR;     // prefetched normals: drivers + one integral per rate
};
using GenericShape = Shape<kMaxRates, kMaxAssets, kMaxCredits, false>;

// Calls f(Shape{}) with the specialisation for (rates, assets, credits), or the generic shape.
template <typename F>
inline auto withShape(int nrates, int nassets, int ncredits, F&& f) {
  switch (nrates * 100 + nassets * 10 + ncredits) {
    case 100: return f(Shape<1, 0, 0, true>{});
    case 101: return f(Shape<1, 0, 1, true>{});
    case 102: return f(Shape<1, 0, 2, true>{});
    case 110: return f(Shape<1, 1, 0, true>{});
    case 111: return f(Shape<1, 1, 1, true>{});
    case 112: return f(Shape<1, 1, 2, true>{});
    case 320: return f(Shape<3, 2, 0, true>{});
    case 321: return f(Shape<3, 2, 1, true>{});
    case 322: return f(Shape<3, 2, 2, true>{});
    default: return f(GenericShape{});
  }
}

// ---------------------------------------------------------------- one path's market
// Risk-factor state of one path (rates, deflator, assets, credits) and the normals of the next
// step, loaded one step ahead: correlated drivers, then the integral normal of each rate
// (domestic right after the drivers, foreign ones at foreign_off).
template <typename Real, typename Sh>
struct Market {
  Real x[Sh::NR];
  Real S[Sh::NA > 0 ? Sh::NA : 1], lnS[Sh::NA > 0 ? Sh::NA : 1], lnSc[Sh::NA > 0 ? Sh::NA : 1];
  Real y[Sh::NC > 0 ? Sh::NC : 1], iy[Sh::NC > 0 ? Sh::NC : 1], Qs[Sh::NC > 0 ? Sh::NC : 1];
  Real lnD, lnDc, D;  // FP32: Kahan sum of int r; FP64: running product
  Real zn[Sh::kPre];
  double prevQ[Sh::NC > 0 ? Sh::NC : 1];  // each credit's survival on the previous reporting date
  double prevD;                            // the deflator on the previous reporting date
};

template <typename Real, typename Sh>
__host__ __device__ __forceinline__ void marketLoad(Market<Real, Sh>& mk, const CCR_CONST GlobalDev& g, const Real* zbase,
                                                    int s) {
  const int nR = Sh::kExact ? Sh::NR : g.nrates, nCorr = Sh::kExact ? Sh::kCorr : g.ncorr;
  const Real* z = zbase + static_cast<int64_t>(s) * g.nz;
#pragma unroll
  for (int i = 0; i < Sh::kCorr; ++i) mk.zn[i] = i < nCorr ? z[i] : Real(0);
#pragma unroll
  for (int r = 0; r < Sh::NR; ++r) mk.zn[Sh::kCorr + r] = r < nR ? z[r == 0 ? nCorr : g.foreign_off + r - 1] : Real(0);
}

// kPrefetch: the normals of step j are loaded during step j - 1 (latency hidden behind one date's
// valuation); without it they are loaded at the step itself (no registers held across dates).
template <typename Real, typename Sh, bool kPrefetch = true>
__host__ __device__ __forceinline__ void marketInit(Market<Real, Sh>& mk, const CCR_CONST GlobalDev& g, const Real* zbase,
                                                    int numDates) {
#pragma unroll
  for (int r = 0; r < Sh::NR; ++r) mk.x[r] = 0;
#pragma unroll
  for (int i = 0; i < Sh::NA; ++i) {
    mk.S[i] = static_cast<Real>(g.spot[i]);
    mk.lnS[i] = 0;
    mk.lnSc[i] = 0;
  }
#pragma unroll
  for (int c = 0; c < Sh::NC; ++c) {
    mk.y[c] = static_cast<Real>(g.cir_y0[c]);
    mk.iy[c] = 0;
    mk.Qs[c] = 1;
    mk.prevQ[c] = 1.0;
  }
  mk.prevD = 1.0;
  mk.lnD = 0;
  mk.lnDc = 0;
  mk.D = 1;
  if (kPrefetch && numDates > 1) marketLoad(mk, g, zbase, 0);
}

// The step t_{j-1} -> t_j (j >= 1) of every risk factor, from the prefetched normals (then the
// next step's are loaded): w = L z (scenario_generator.cpp:118-124), exact Hull-White steps
// (hull_white.cpp:40-68; foreign rates with the quanto drift), deflator D *= exp(-int r)
// (scenario_generator.cpp:126-128), lognormal assets (lognormal_asset.cpp:17-22) with the carry of
// a curve or a foreign rate, CIR++ bridge sub-steps and pathwise survival (cir_intensity.cpp:45-66).
template <typename Real, typename Sh, bool kPrefetch = true>
__host__ __device__ __forceinline__ void marketStep(Market<Real, Sh>& mk, const CCR_CONST GlobalDev& g, int j, int numDates,
                                                    cptr<Real> steps, const Real* zbase, Real sgn) {
  constexpr bool kDouble = std::is_same<Real, double>::value;
  constexpr int NR = Sh::NR, NA = Sh::NA, NC = Sh::NC;
  const int nR = Sh::kExact ? NR : g.nrates, nA = Sh::kExact ? NA : g.nassets, nC = Sh::kExact ? NC : g.ncredits;
  cptr<Real> st = steps + static_cast<int64_t>(j) * kStepStride;
  const Real* z = zbase + static_cast<int64_t>(j - 1) * g.nz;
  Real zc[Sh::kPre];
  if (!kPrefetch) marketLoad(mk, g, zbase, j - 1);
#pragma unroll
  for (int i = 0; i < Sh::kPre; ++i) zc[i] = sgn * mk.zn[i];
  if (kPrefetch && j + 1 < numDates) marketLoad(mk, g, zbase, j);
  // Correlated drivers at fixed factor slots: rates [0, kMaxRates), assets, credits. Row `slot`
  // of the factor holds compact row `lim` (lower triangular: columns k <= lim).
  Real w[Sh::kCorr];
  {
    cptr<double> chol = g.chol;
#if defined(__HIP_DEVICE_COMPILE__)
    // Opaque per step: the factors are re-read through the scalar cache each step instead of
    // being hoisted out of the date loop into registers.
    asm volatile("" : "+s"(chol));
#endif
#pragma unroll
    for (int i = 0; i < Sh::kCorr; ++i) {
      const int slot = i < NR ? i : i < NR + NA ? kMaxRates + (i - NR) : kMaxRates + kMaxAssets + (i - NR - NA);
      const int lim = i < NR ? i : i < NR + NA ? nR + (i - NR) : nR + nA + (i - NR - NA);
      const bool used = i < NR ? i < nR : i < NR + NA ? i - NR < nA : i - NR - NA < nC;
      Real sum = 0;
      if (used) {
#pragma unroll
        for (int k = 0; k < Sh::kCorr; ++k)
          if (k <= lim) sum += static_cast<Real>(chol[slot * kMaxCorr + k]) * zc[k];
      }
      w[i] = sum;
    }
  }
  Real ir[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r < nR) {
      cptr<Real> o = st + kStepRate + 8 * r;
      const Real z2 = zc[Sh::kCorr + r];
      const Real x0 = mk.x[r];
      mk.x[r] = x0 * o[0] + o[1] * w[r] + o[5];
      const Real ix = x0 * o[4] + o[2] * w[r] + o[3] * z2 + o[6];
      ir[r] = ix + o[7];
    } else {
      ir[r] = 0;
    }
  }
  if constexpr (kDouble) {
    mk.D *= exp(-ir[0]);
  } else {
    const Real yk = ir[0] - mk.lnDc;  // Kahan-compensated integral of the short rate
    const Real tk = mk.lnD + yk;
    mk.lnDc = (tk - mk.lnD) - yk;
    mk.lnD = tk;
    mk.D = expr(-mk.lnD);
  }
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    if (i < nA) {
      cptr<Real> o = st + kStepAsset + 3 * i;
      const int rr = g.asset_rate[i];
      Real icarry = o[0];
#pragma unroll
      for (int r = 1; r < NR; ++r)
        if (r == rr) icarry = ir[r];
      const Real drift = ir[0] - icarry - o[2] + o[1] * w[NR + i];
      if constexpr (kDouble) {
        mk.S[i] = mk.S[i] * exp(drift);
      } else {
        const Real yk = drift - mk.lnSc[i];
        const Real tk = mk.lnS[i] + yk;
        mk.lnSc[i] = (tk - mk.lnS[i]) - yk;
        mk.lnS[i] = tk;
        mk.S[i] = static_cast<Real>(g.spot[i]) * expr(mk.lnS[i]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c < nC) {
      cptr<Real> o = st + kStepCredit + kStepCreditStride * c;
      const int sub = g.credit_sub[c];
      const Real* ex = z + g.credit_off[c];
      const Real kappa = static_cast<Real>(g.cir_kappa[c]), theta = static_cast<Real>(g.cir_theta[c]);
      const Real xi = static_cast<Real>(g.cir_xi[c]), h = o[1];
      Real remaining = o[0] * w[NR + NA + c];
      Real yc = mk.y[c], iyc = mk.iy[c];
      for (int k = 0; k < sub; ++k) {
        Real dW = remaining;
        if (k + 1 < sub) dW = o[3 + k] * remaining + o[3 + kMaxSub + k] * (sgn * ex[k]);
        remaining -= dW;
        const Real yp = yc > Real(0) ? yc : Real(0);
        const Real yn = yc + kappa * (theta - yp) * h + xi * sqrtr(yp) * dW;
        iyc += Real(0.5) * (yp + (yn > Real(0) ? yn : Real(0))) * h;
        yc = yn;
      }
      mk.y[c] = yc;
      mk.iy[c] = iyc;
      mk.Qs[c] = expr(-iyc - o[2]);
    }
  }
}

// ---------------------------------------------------------------- nested initial margin
// Inner paths of the nested conditional initial margin draw their normals from a counter-based
// generator: SplitMix64's finaliser (Steele, Lea and Flood, 2014) as the mixing function, keys
// chained over (seed, netting set, outer path, reporting date, antithetic pair), and normal k of
// inner step j (k the index in a step's normal layout: correlated drivers, then the domestic
// integral normal, then the foreign ones) by the Box-Muller cosine branch of two uniforms from
// mix64(c) and mix64(c ^ constant), c = key ^ (j << 16 | k).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t nestedPathKey(uint64_t seed, int set, int path) {
  return mix64(mix64(mix64(seed) ^ static_cast<uint64_t>(set)) ^ static_cast<uint64_t>(path));
}
__host__ __device__ inline uint64_t nestedPairKey(uint64_t pathKey, int date, int pair) {
  return mix64(mix64(pathKey ^ (static_cast<uint64_t>(date) << 32)) ^ static_cast<uint64_t>(pair));
}
__host__ __device__ inline double nestedNormal(uint64_t key, int step, int k) {
  const uint64_t c = key ^ ((static_cast<uint64_t>(step) << 16) | static_cast<uint64_t>(k));
  const uint64_t a = mix64(c), b = mix64(c ^ 0xD1B54A32D192ED03ull);
  const double u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// One inner step t_{j-1} -> t_j of the rates and assets (the credit factors do not enter a
// netting set's value and are not stepped): marketStep's arithmetic on normals from the inner
// generator (key, sign: the antithetic partner flips every normal).
template <typename Real, typename Sh>
__host__ __device__ __forceinline__ void innerStep(Market<Real, Sh>& mk, const CCR_CONST GlobalDev& g, int j,
                                                   cptr<Real> steps, uint64_t key, double sign) {
  constexpr bool kDouble = std::is_same<Real, double>::value;
  constexpr int NR = Sh::NR, NA = Sh::NA;
  const int nR = Sh::kExact ? NR : g.nrates, nA = Sh::kExact ? NA : g.nassets;
  const int nCorr = Sh::kExact ? Sh::kCorr : g.ncorr;
  cptr<Real> st = steps + static_cast<int64_t>(j) * kStepStride;
  Real zc[NR + NA > 0 ? NR + NA : 1];
#pragma unroll
  for (int k = 0; k < NR + NA; ++k) {
    const bool used = k < NR ? k < nR : k - NR < nA;
    const int idx = k < NR ? k : nR + (k - NR);  // compact driver index
    zc[k] = used ? static_cast<Real>(sign * nestedNormal(key, j, idx)) : Real(0);
  }
  Real z2[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r)
    z2[r] = r < nR ? static_cast<Real>(sign * nestedNormal(key, j, r == 0 ? nCorr : g.foreign_off + r - 1)) : Real(0);
  // Correlated drivers of the rates and assets (rows of the factor with columns k <= lim, compact).
  Real w[NR + NA > 0 ? NR + NA : 1];
  cptr<double> chol = g.chol;
#pragma unroll
  for (int i = 0; i < NR + NA; ++i) {
    const int slot = i < NR ? i : kMaxRates + (i - NR);
    const int lim = i < NR ? i : nR + (i - NR);
    const bool used = i < NR ? i < nR : i - NR < nA;
    Real sum = 0;
    if (used) {
#pragma unroll
      for (int k = 0; k < NR + NA; ++k) {
        const int ck = k < NR ? k : nR + (k - NR);  // compact index of slot k
        const bool kused = k < NR ? k < nR : k - NR < nA;
        if (kused && ck <= lim) sum += static_cast<Real>(chol[slot * kMaxCorr + ck]) * zc[k];
      }
    }
    w[i] = sum;
  }
  Real ir[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r < nR) {
      cptr<Real> o = st + kStepRate + 8 * r;
      const Real x0 = mk.x[r];
      mk.x[r] = x0 * o[0] + o[1] * w[r] + o[5];
      const Real ix = x0 * o[4] + o[2] * w[r] + o[3] * z2[r] + o[6];
      ir[r] = ix + o[7];
    } else {
      ir[r] = 0;
    }
  }
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    if (i < nA) {
      cptr<Real> o = st + kStepAsset + 3 * i;
      const int rr = g.asset_rate[i];
      Real icarry = o[0];
#pragma unroll
      for (int r = 1; r < NR; ++r)
        if (r == rr) icarry = ir[r];
      const Real drift = ir[0] - icarry - o[2] + o[1] * w[NR + i];
      if constexpr (kDouble) {
        mk.S[i] = mk.S[i] * exp(drift);
      } else {
        const Real yk = drift - mk.lnSc[i];
        const Real tk = mk.lnS[i] + yk;
        mk.lnSc[i] = (tk - mk.lnS[i]) - yk;
        mk.lnS[i] = tk;
        mk.S[i] = static_cast<Real>(g.spot[i]) * expr(mk.lnS[i]);
      }
    }
  }
}

// Fixing states recorded on date j: x of each listed rate into its slot row.
// One predicated store per rate with a constant index into the state: a select among the rates'
// states would be folded into a dynamically indexed load, which keeps the whole market state in
// scratch memory.
template <typename Real, typename Sh>
__host__ __device__ __forceinline__ void storeFixings(const Market<Real, Sh>& mk, int beg, int end, cptr<int32_t> fstore,
                                                      Real* xs, int stride) {
  for (int i = beg; i < end; ++i) {
    const int r = fstore[2 * i], slot = fstore[2 * i + 1];
#pragma unroll
    for (int k = 0; k < Sh::NR; ++k)
      if (k == r) xs[slot * stride] = mk.x[k];
  }
}

// Value of the terms of one valuation record on its date: V = sum over rate groups of FX x (bond
// polynomial or merged bond terms + fixed coupons) + deterministic-carry forwards (+ options).
template <typename Real, typename Sh, bool kOptions, typename Rec>
__host__ __device__ __forceinline__ Real valueRecord(const Market<Real, Sh>& mk, const CCR_CONST GlobalDev& g,
                                                     const Rec& vr, cptr<Real> vdr, cptr<Real> poly,
                                                     cptr<Real> bonds, cptr<Real> fixed, cptr<Real> opts,
                                                     const Real* xs, int stride) {
  constexpr int NR = Sh::NR, NA = Sh::NA;
  const int nR = Sh::kExact ? NR : g.nrates, nA = Sh::kExact ? NA : g.nassets;
  Real Vt = 0;
  // Merged bonds (and their constants) of every rate as power series in its state: Horner
  // in FP64 over the interleaved coefficient rows (highest power first, one coefficient per
  // rate per row), the rates' chains side by side and kPolyBatch rows per batch of scalar loads.
  // The coefficients are uniform: the date loop is uniform across the wavefront.
  const int rows = vr.poly_rows;
  bool usePoly = rows > 0;
  if (usePoly) {
    bool inside = true;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < nR) inside = inside && static_cast<double>(mk.x[r] < Real(0) ? -mk.x[r] : mk.x[r]) <= vdr[kDateXmax + r];
    usePoly = allLanes(inside);
  }
  Real hp[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) hp[r] = 0;
  if (usePoly) {
    cptr<Real> pc = poly + vr.poly_beg;
    for (int k = 0; k < rows; k += kPolyBatch) {
      Real c[kPolyBatch][NR];
#pragma unroll
      for (int u = 0; u < kPolyBatch; ++u)
#pragma unroll
        for (int r = 0; r < NR; ++r) c[u][r] = r < nR ? pc[(k + u) * nR + r] : Real(0);
#pragma unroll
      for (int u = 0; u < kPolyBatch; ++u)
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (r < nR) hp[r] = hornerStep<Real>(hp[r], mk.x[r], c[u][r]);
    }
  }
  // Fixed floating coupons of every rate (contiguous over the rates), two per batch so that
  // their record loads, fixing-state reads and exponentials overlap; the plan pads the coupon
  // table with one zero record, so the second record of a batch always exists.
  Real fsum[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) fsum[r] = 0;
  {
    const int fe = vr.fix_beg[kMaxRates];
    for (int i = vr.fix_beg[0]; i < fe; i += 2) {
      cptr<Real> f = fixed + 4 * static_cast<int64_t>(i);
      const Real a0 = f[0], b0 = f[1], c0 = f[2], a1 = f[4], b1 = f[5], c1 = f[6];
      const int s0 = static_cast<int>(f[3]), s1 = static_cast<int>(f[7]);
      const bool two = i + 1 < fe;
      int r0 = 0, r1 = 0;
#pragma unroll
      for (int r = 1; r < NR; ++r) {
        if (i >= vr.fix_beg[r]) r0 = r;
        if (i + 1 >= vr.fix_beg[r]) r1 = r;
      }
      Real x0 = mk.x[0], x1 = mk.x[0];
#pragma unroll
      for (int r = 1; r < NR; ++r) {
        if (r == r0) x0 = mk.x[r];
        if (r == r1) x1 = mk.x[r];
      }
      const Real xk0 = xs[s0 * stride], xk1 = xs[s1 * stride];
      const Real t0 = a0 * expTerm(-b0 * x0 - c0 * xk0);
      const Real t1 = a1 * expTerm(-b1 * x1 - c1 * xk1);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r == r0) fsum[r] += t0;
        if (two && r == r1) fsum[r] += t1;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (r < nR) {
      const Real nx = -mk.x[r];
      Real acc;
      if (usePoly) {
        acc = hp[r];
      } else {
        acc = vdr[kDateC0 + r];
        for (int i = vr.bond_beg[r]; i < vr.bond_beg[r + 1]; ++i) acc += bonds[2 * i] * expTerm(bonds[2 * i + 1] * nx);
      }
      acc += fsum[r];
      Real mult = 1;
      if (r > 0) {
        const int fa = g.fx_of[r];
#pragma unroll
        for (int i = 0; i < NA; ++i)
          if (i == fa) mult = mk.S[i];
      }
      Vt += mult * acc;
    }
  }
#pragma unroll
  for (int i = 0; i < NA; ++i)
    if (i < nA) Vt += mk.S[i] * vdr[kDateLin + i];
  if constexpr (NA > 0 && kOptions) {
    for (int i = vr.opt_beg; i < vr.opt_end; ++i) {
      cptr<Real> o = opts + kOptStride * static_cast<int64_t>(i);
      const int oa = static_cast<int>(o[0]);
      Real Sa = mk.S[0];
#pragma unroll
      for (int k = 1; k < NA; ++k)
        if (k == oa) Sa = mk.S[k];
      const Real df = o[4] * expr(-o[5] * mk.x[0]);
      Vt += o[2] * blackFormula(o[1], Sa * o[6] / df, o[3], o[7], df);
    }
  }
  return Vt;
}

// Per-path state a netting set carries from date to date: CSA collateral units and the previous
// reporting date's V - C, whose positive and negative parts times that date's deflator
// (Market::prevD) are its discounted exposures (pathwise CVA / DVA trapezoids; the survivals of
// that date are the market's, Market::prevQ). One double per set instead of two: the deflator
// is shared by the sets of a lane.
template <typename Real>
struct SetPath {
  Real units;
  double prevV;
};
template <typename Real>
__host__ __device__ __forceinline__ void setPathInit(SetPath<Real>& st) {
  st.units = 0;
  st.prevV = 0;
}

// After the sets' reports of a reporting date: its survivals and deflator become the previous
// ones.
template <typename Real, typename Sh>
__host__ __device__ __forceinline__ void marketReported(Market<Real, Sh>& mk) {
#pragma unroll
  for (int c = 0; c < Sh::NC; ++c) mk.prevQ[c] = static_cast<double>(mk.Qs[c]);
  mk.prevD = static_cast<double>(mk.D);
}

// Collateral (collateral.cpp:10-47, the margin-call value read from the set's ring) and, on a
// reporting date, the 9 statistics of ExposureProfile::compute and pathwiseCva for a set whose
// value before collateral on date j is V; emit(rep, stats[kStats], positive_exposure).
template <typename Real, typename Sh, typename Rec, typename Emit>
__host__ __device__ __forceinline__ Real collateralAndReport(const Market<Real, Sh>& mk, const CCR_CONST SetDev& S,
                                                            const Rec& rec, cptr<Real> dr, int j, Real V,
                                                            SetPath<Real>& st, Real* ringV, Real* ringP, int stride,
                                                            Emit&& emit) {
  // Not contracted into FMAs: the per-set kernel's copy of this arithmetic rounds alike.
#pragma clang fp contract(off)
  constexpr int NA = Sh::NA, NC = Sh::NC;
  Real C = 0;
  if (S.has_csa) {
    Real price = 1;
    if (S.sec_type == CCR_COLLATERAL_BOND) {
      // expTerm's FP64 polynomial reads its coefficients from SGPRs: the library exp keeps them
      // in VGPRs, which the book kernel's unrolled set loop hoists and spills for every set.
      price = dr[kDateBond] * expTerm(-dr[kDateBond + 1] * mk.x[0]);
    } else if (S.sec_type == CCR_COLLATERAL_ASSET) {
#pragma unroll
      for (int i = 0; i < NA; ++i)
        if (i == S.csa_asset) price = mk.S[i];
    }
    const bool sec = S.sec_type != CCR_COLLATERAL_CASH;
    if (rec.store_pos >= 0) {
      ringV[rec.store_pos * stride] = V;
      if (sec) ringP[rec.store_pos * stride] = price;
    }
    const Real vc = ringV[rec.call_pos * stride];
    const Real pc = sec ? ringP[rec.call_pos * stride] : Real(1);
    const Real keep = static_cast<Real>(S.keep), thc = static_cast<Real>(S.thr_cpty);
    const Real tho = static_cast<Real>(S.thr_own);
    Real target = 0;
    if (vc > thc) target += vc - thc;
    if (S.thr_own >= 0.0 && -vc > tho) target -= -vc - tho;
    if (j == 0) st.units = target / (keep * pc);
    Real diff = target - st.units * keep * pc;
    diff = diff < Real(0) ? -diff : diff;
    if (diff >= static_cast<Real>(S.mta)) st.units = target / (keep * pc);
    C = st.units * price + static_cast<Real>(S.ia);
  }
  if (!(rec.flags & kReporting)) return C;
  const double v = static_cast<double>(V) - static_cast<double>(C);
  const double d = static_cast<double>(mk.D);
  const double vp = v > 0.0 ? v : 0.0, vn = v < 0.0 ? v : 0.0;
  const double dvp = d * vp, dvn = d * (-vn);
  double qc = 1.0, qo = 1.0, pqc = 1.0, pqo = 1.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c == S.cpty) {
      qc = static_cast<double>(mk.Qs[c]);
      pqc = mk.prevQ[c];
    }
    if (c == S.own) {
      qo = static_cast<double>(mk.Qs[c]);
      pqo = mk.prevQ[c];
    }
  }
  const int rr = rec.rep;
  double st9[kStats];
  st9[0] = v;
  st9[1] = vp;
  st9[2] = vn;
  st9[3] = dvp;
  st9[4] = d * vn;
  st9[5] = d;
  st9[6] = qc;
  // The previous date's discounted exposures, recomputed from its V - C and deflator by the same
  // operations that formed dvp and dvn then (so bit for bit the same values).
  const double pv = st.prevV;
  const double pvp = pv > 0.0 ? pv : 0.0, pvn = pv < 0.0 ? pv : 0.0;
  const double pdvp = mk.prevD * pvp, pdvn = mk.prevD * (-pvn);
  st9[7] = rr > 0 ? 0.5 * (pdvp + dvp) * (pqc - qc) : 0.0;
  st9[8] = rr > 0 ? 0.5 * (pdvn + dvn) * (pqo - qo) : 0.0;
  st.prevV = v;
  emit(rr, st9, vp);
  return C;
}

// The per-set kernel's body keeps its own copy of the market step and valuation above (inlined
// by hand): its register allocation is tuned to five waves per SIMD and the shared helpers, used
// by the book kernel below, change it.
// col: this path's column of per-path scratch (LDS on the GPU), stride: distance between rows;
// rows [0, ring) hold the margin-call values, [ring, ring + priceRows) the collateral prices,
// then the fixing states. emit(rep, stats[kStats], positive_exposure) is called on every
// reporting date.
// kOptions: the plan holds European options (the Black evaluation is compiled in only then, so
// option-free plans keep a compact loop). onValue(j, V, D) is called on every valued date with
// the netted value before collateral and the deflator.
// kTrades (per-trade plans): each of the set's S.num_groups trades is valued from its own records
// (tdates / tdreals, [j G + i]) into row i of tv, the netted value is their sum in trade order,
// and onTrades(rep, V, D, tv, stride) is called on every reporting date.
template <typename Real, typename Sh, bool kOptions, bool kTrades, typename Emit, typename OnValue,
          typename OnTrades>
__host__ __device__ inline void simulatePath(cptr<GlobalDev> gp, cptr<SetDev> sp, int pe,
                                             const Real* __restrict__ zbase, cptr<Real> steps, cptr<DateRec> dates,
                                             cptr<Real> dreal, cptr<Real> bonds, cptr<Real> fixed,
                                             cptr<int32_t> fstore, cptr<Real> poly, cptr<Real> opts,
                                             const double* bermv, int ring,
                                             int priceRows, Real* col, int stride, Emit&& emit,
                                             OnValue&& onValue, cptr<DateRec> tdates, cptr<Real> tdreals,
                                             Real* tv, OnTrades&& onTrades) {
  constexpr bool kDouble = std::is_same<Real, double>::value;
  constexpr int NR = Sh::NR, NA = Sh::NA, NC = Sh::NC;
  const CCR_CONST GlobalDev& g = *gp;
  const CCR_CONST SetDev& S = *sp;
  Real* ringV = col;
  Real* ringP = col + ring * stride;
  Real* xs = col + (ring + priceRows) * stride;
  const int nR = Sh::kExact ? NR : g.nrates, nA = Sh::kExact ? NA : g.nassets, nC = Sh::kExact ? NC : g.ncredits;
  const int nz = g.nz, nCorr = Sh::kExact ? Sh::kCorr : g.ncorr;
  const Real sgn = (g.antithetic && (pe & 1)) ? Real(-1) : Real(1);

  Real x[NR], S_[NA > 0 ? NA : 1], y[NC > 0 ? NC : 1], iy[NC > 0 ? NC : 1], Qs[NC > 0 ? NC : 1];
  Real lnD = 0, lnDc = 0, D = 1;  // FP32: Kahan sum of int r; FP64: running product
  Real lnS[NA > 0 ? NA : 1], lnSc[NA > 0 ? NA : 1];
#pragma unroll
  for (int r = 0; r < NR; ++r) x[r] = 0;
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    S_[i] = static_cast<Real>(g.spot[i]);
    lnS[i] = 0;
    lnSc[i] = 0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    y[c] = static_cast<Real>(g.cir_y0[c]);
    iy[c] = 0;
    Qs[c] = 1;
  }
  Real units = 0;
  const double* bermRow = bermv ? bermv + static_cast<int64_t>(S.cd_base) * g.num_paths + pe : nullptr;
  double prevDvp = 0, prevDvn = 0, prevQc = 1, prevQo = 1;

  // Normals of the next step, loaded one step ahead: correlated drivers, then the integral
  // normal of each rate (domestic right after the drivers, foreign ones at foreign_off).
  Real zn[Sh::kPre];
  auto loadStep = [&](int s) {
    const Real* z = zbase + static_cast<int64_t>(s) * nz;
#pragma unroll
    for (int i = 0; i < Sh::kCorr; ++i) zn[i] = i < nCorr ? z[i] : Real(0);
#pragma unroll
    for (int r = 0; r < NR; ++r) zn[Sh::kCorr + r] = r < nR ? z[r == 0 ? nCorr : g.foreign_off + r - 1] : Real(0);
  };
  if (S.num_dates > 1) loadStep(0);

  for (int j = 0; j < S.num_dates; ++j) {
    if (j > 0) {
      cptr<Real> st = steps + static_cast<int64_t>(j) * kStepStride;
      const Real* z = zbase + static_cast<int64_t>(j - 1) * nz;
      Real zc[Sh::kPre];
#pragma unroll
      for (int i = 0; i < Sh::kPre; ++i) zc[i] = sgn * zn[i];
      if (j + 1 < S.num_dates) loadStep(j);
      // Correlated drivers at fixed factor slots: rates [0, kMaxRates), assets, credits. Row
      // `slot` of the factor holds compact row `lim` (lower triangular: columns k <= lim).
      Real w[Sh::kCorr];
      {
        cptr<double> chol = g.chol;
#if defined(__HIP_DEVICE_COMPILE__)
        // Opaque per step: the factors are re-read through the scalar cache each step instead of
        // being hoisted out of the date loop into registers.
        asm volatile("" : "+s"(chol));
#endif
#pragma unroll
        for (int i = 0; i < Sh::kCorr; ++i) {
          const int slot = i < NR ? i : i < NR + NA ? kMaxRates + (i - NR) : kMaxRates + kMaxAssets + (i - NR - NA);
          const int lim = i < NR ? i : i < NR + NA ? nR + (i - NR) : nR + nA + (i - NR - NA);
          const bool used = i < NR ? i < nR : i < NR + NA ? i - NR < nA : i - NR - NA < nC;
          Real sum = 0;
          if (used) {
#pragma unroll
            for (int k = 0; k < Sh::kCorr; ++k)
              if (k <= lim) sum += static_cast<Real>(chol[slot * kMaxCorr + k]) * zc[k];
          }
          w[i] = sum;
        }
      }
      Real ir[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r < nR) {
          cptr<Real> o = st + kStepRate + 8 * r;
          const Real z2 = zc[Sh::kCorr + r];
          const Real x0 = x[r];
          x[r] = x0 * o[0] + o[1] * w[r] + o[5];
          const Real ix = x0 * o[4] + o[2] * w[r] + o[3] * z2 + o[6];
          ir[r] = ix + o[7];
        } else {
          ir[r] = 0;
        }
      }
      if constexpr (kDouble) {
        D *= exp(-ir[0]);
      } else {
        const Real yk = ir[0] - lnDc;  // Kahan-compensated integral of the short rate
        const Real tk = lnD + yk;
        lnDc = (tk - lnD) - yk;
        lnD = tk;
        D = expr(-lnD);
      }
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        if (i < nA) {
          cptr<Real> o = st + kStepAsset + 3 * i;
          const int rr = g.asset_rate[i];
          Real icarry = o[0];
#pragma unroll
          for (int r = 1; r < NR; ++r)
            if (r == rr) icarry = ir[r];
          const Real drift = ir[0] - icarry - o[2] + o[1] * w[NR + i];
          if constexpr (kDouble) {
            S_[i] = S_[i] * exp(drift);
          } else {
            const Real yk = drift - lnSc[i];
            const Real tk = lnS[i] + yk;
            lnSc[i] = (tk - lnS[i]) - yk;
            lnS[i] = tk;
            S_[i] = static_cast<Real>(g.spot[i]) * expr(lnS[i]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c < nC) {
          cptr<Real> o = st + kStepCredit + kStepCreditStride * c;
          const int sub = g.credit_sub[c];
          const Real* ex = z + g.credit_off[c];
          const Real kappa = static_cast<Real>(g.cir_kappa[c]), theta = static_cast<Real>(g.cir_theta[c]);
          const Real xi = static_cast<Real>(g.cir_xi[c]), h = o[1];
          Real remaining = o[0] * w[NR + NA + c];
          Real yc = y[c], iyc = iy[c];
          for (int k = 0; k < sub; ++k) {
            Real dW = remaining;
            if (k + 1 < sub) dW = o[3 + k] * remaining + o[3 + kMaxSub + k] * (sgn * ex[k]);
            remaining -= dW;
            const Real yp = yc > Real(0) ? yc : Real(0);
            const Real yn = yc + kappa * (theta - yp) * h + xi * sqrtr(yp) * dW;
            iyc += Real(0.5) * (yp + (yn > Real(0) ? yn : Real(0))) * h;
            yc = yn;
          }
          y[c] = yc;
          iy[c] = iyc;
          Qs[c] = expr(-iyc - o[2]);
        }
      }
    }

    const DateRec rec = dates[j];
    for (int i = rec.fixstore_beg; i < rec.fixstore_end; ++i) {
      const int r = fstore[2 * i], slot = fstore[2 * i + 1];
#pragma unroll
      for (int k = 0; k < NR; ++k)
        if (k == r) xs[slot * stride] = x[k];  // constant-index predicated stores (see storeFixings)
    }
    if (!(rec.flags & kValued) && !S.has_csa) continue;
    cptr<Real> dr = dreal + static_cast<int64_t>(j) * kDateRealStride;
    // Value of the terms of one valuation record on this date.
    auto valueTerms = [&](const DateRec& vr, cptr<Real> vdr) -> Real {
      Real Vt = 0;
      // Merged bonds (and their constants) of every rate as power series in its state: Horner
      // in FP64 over the interleaved coefficient rows (highest power first, one coefficient per
      // rate per row), the rates' chains side by side and kPolyBatch rows per batch of scalar loads.
      // The coefficients are uniform: the date loop is uniform across the wavefront.
      const int rows = vr.poly_rows;
      bool usePoly = rows > 0;
      if (usePoly) {
        bool inside = true;
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (r < nR) inside = inside && static_cast<double>(x[r] < Real(0) ? -x[r] : x[r]) <= vdr[kDateXmax + r];
        usePoly = allLanes(inside);
      }
      Real hp[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) hp[r] = 0;
      if (usePoly) {
        cptr<Real> pc = poly + vr.poly_beg;
        for (int k = 0; k < rows; k += kPolyBatch) {
          Real c[kPolyBatch][NR];
#pragma unroll
          for (int u = 0; u < kPolyBatch; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r) c[u][r] = r < nR ? pc[(k + u) * nR + r] : Real(0);
#pragma unroll
          for (int u = 0; u < kPolyBatch; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r)
              if (r < nR) hp[r] = hornerStep<Real>(hp[r], x[r], c[u][r]);
        }
      }
      // Fixed floating coupons of every rate (contiguous over the rates), two per batch so that
      // their record loads, fixing-state reads and exponentials overlap; the plan pads the coupon
      // table with one zero record, so the second record of a batch always exists.
      Real fsum[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) fsum[r] = 0;
      {
        const int fe = vr.fix_beg[kMaxRates];
        for (int i = vr.fix_beg[0]; i < fe; i += 2) {
          cptr<Real> f = fixed + 4 * static_cast<int64_t>(i);
          const Real a0 = f[0], b0 = f[1], c0 = f[2], a1 = f[4], b1 = f[5], c1 = f[6];
          const int s0 = static_cast<int>(f[3]), s1 = static_cast<int>(f[7]);
          const bool two = i + 1 < fe;
          int r0 = 0, r1 = 0;
#pragma unroll
          for (int r = 1; r < NR; ++r) {
            if (i >= vr.fix_beg[r]) r0 = r;
            if (i + 1 >= vr.fix_beg[r]) r1 = r;
          }
          Real x0 = x[0], x1 = x[0];
#pragma unroll
          for (int r = 1; r < NR; ++r) {
            if (r == r0) x0 = x[r];
            if (r == r1) x1 = x[r];
          }
          const Real xk0 = xs[s0 * stride], xk1 = xs[s1 * stride];
          const Real t0 = a0 * expTerm(-b0 * x0 - c0 * xk0);
          const Real t1 = a1 * expTerm(-b1 * x1 - c1 * xk1);
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            if (r == r0) fsum[r] += t0;
            if (two && r == r1) fsum[r] += t1;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r < nR) {
          const Real nx = -x[r];
          Real acc;
          if (usePoly) {
            acc = hp[r];
          } else {
            acc = vdr[kDateC0 + r];
            for (int i = vr.bond_beg[r]; i < vr.bond_beg[r + 1]; ++i) acc += bonds[2 * i] * expTerm(bonds[2 * i + 1] * nx);
          }
          acc += fsum[r];
          Real mult = 1;
          if (r > 0) {
            const int fa = g.fx_of[r];
#pragma unroll
            for (int i = 0; i < NA; ++i)
              if (i == fa) mult = S_[i];
          }
          Vt += mult * acc;
        }
      }
#pragma unroll
      for (int i = 0; i < NA; ++i)
        if (i < nA) Vt += S_[i] * vdr[kDateLin + i];
      if constexpr (NA > 0 && kOptions) {
        for (int i = vr.opt_beg; i < vr.opt_end; ++i) {
          cptr<Real> o = opts + kOptStride * static_cast<int64_t>(i);
          const int oa = static_cast<int>(o[0]);
          Real Sa = S_[0];
#pragma unroll
          for (int k = 1; k < NA; ++k)
            if (k == oa) Sa = S_[k];
          const Real df = o[4] * expr(-o[5] * x[0]);
          Vt += o[2] * blackFormula(o[1], Sa * o[6] / df, o[3], o[7], df);
        }
      }
      return Vt;
    };
    Real V = 0;
    if (rec.flags & kValued) {
      V = valueTerms(rec, dr);
      if constexpr (kTrades) {
        const int G = S.num_groups;
        cptr<DateRec> trecs = tdates + static_cast<int64_t>(j) * G;
        cptr<Real> tdr = tdreals + static_cast<int64_t>(j) * G * kDateRealStride;
        for (int i = 0; i < G; ++i) {
          const DateRec tr = trecs[i];
          const Real vi = valueTerms(tr, tdr + i * kDateRealStride);
          tv[i * stride] = vi;
          V += vi;
        }
      }
      // Bermudan swaptions: their per-path sum on this valued date, from the valuation kernel
      // (one row per valued date of the set, in date order).
      if (bermRow) {
        V += static_cast<Real>(*bermRow);
        bermRow += g.num_paths;
      }
    }
    Real C = 0;
    if (S.has_csa) {
      // The collateral and statistics arithmetic is not contracted into FMAs (here and in
      // collateralAndReport): every kernel that reports a set rounds them alike.
#pragma clang fp contract(off)
      Real price = 1;
      if (S.sec_type == CCR_COLLATERAL_BOND) {
        price = dr[kDateBond] * expTerm(-dr[kDateBond + 1] * x[0]);
      } else if (S.sec_type == CCR_COLLATERAL_ASSET) {
#pragma unroll
        for (int i = 0; i < NA; ++i)
          if (i == S.csa_asset) price = S_[i];
      }
      const bool sec = S.sec_type != CCR_COLLATERAL_CASH;
      if (rec.store_pos >= 0) {
        ringV[rec.store_pos * stride] = V;
        if (sec) ringP[rec.store_pos * stride] = price;
      }
      const Real vc = ringV[rec.call_pos * stride];
      const Real pc = sec ? ringP[rec.call_pos * stride] : Real(1);
      const Real keep = static_cast<Real>(S.keep), thc = static_cast<Real>(S.thr_cpty);
      const Real tho = static_cast<Real>(S.thr_own);
      Real target = 0;
      if (vc > thc) target += vc - thc;
      if (S.thr_own >= 0.0 && -vc > tho) target -= -vc - tho;
      if (j == 0) units = target / (keep * pc);
      Real diff = target - units * keep * pc;
      diff = diff < Real(0) ? -diff : diff;
      if (diff >= static_cast<Real>(S.mta)) units = target / (keep * pc);
      C = units * price + static_cast<Real>(S.ia);
    }
    if (rec.flags & kValued) onValue(j, V, D);
    if (!(rec.flags & kReporting)) continue;
    {
#pragma clang fp contract(off)
      const double v = static_cast<double>(V) - static_cast<double>(C);
      const double d = static_cast<double>(D);
      const double vp = v > 0.0 ? v : 0.0, vn = v < 0.0 ? v : 0.0;
      const double dvp = d * vp, dvn = d * (-vn);
      double qc = 1.0, qo = 1.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c == S.cpty) qc = static_cast<double>(Qs[c]);
        if (c == S.own) qo = static_cast<double>(Qs[c]);
      }
      const int rr = rec.rep;
      double st9[kStats];
      st9[0] = v;
      st9[1] = vp;
      st9[2] = vn;
      st9[3] = dvp;
      st9[4] = d * vn;
      st9[5] = d;
      st9[6] = qc;
      st9[7] = rr > 0 ? 0.5 * (prevDvp + dvp) * (prevQc - qc) : 0.0;
      st9[8] = rr > 0 ? 0.5 * (prevDvn + dvn) * (prevQo - qo) : 0.0;
      prevDvp = dvp;
      prevDvn = dvn;
      prevQc = qc;
      prevQo = qo;
      emit(rr, st9, vp);
      if constexpr (kTrades) onTrades(rr, V, D, tv, stride);
    }
  }
}

// Book plans, G netting sets per lane: the sets of a book share their grid and their draws, so
// one lane steps its path's market once per date and values every set of its group on it. Per
// path, LDS holds the shared fixing states (rows [0, slots), recorded from the group's first
// set's lists -- a book's lists are the same) and each set's margin-call ring and price rows (rows
// slots + k (ring + priceRows) ...). Each set's collateral units and previous discounted
// exposures stay in registers when the set loop is unrolled (groupUnrolled(G)); otherwise the
// loop stays rolled and they live in `state` (a double column, rows 2k and 2k + 1), since
// registers indexed by a loop counter would sit in scratch memory.
// sp[g], g < ns: the group's sets; emit(g, rep, stats, positive_exposure).
template <typename Real, typename Sh, int G, typename Emit>
__host__ __device__ __forceinline__ void simulateGroup(cptr<GlobalDev> gp, cptr<SetDev> sets, int s0, int ns, int pe,
                                              const Real* __restrict__ zbase, cptr<Real> steps, cptr<DateRec> dates,
                                              cptr<Real> dreal, cptr<Real> bonds, cptr<Real> fixed,
                                              cptr<int32_t> fstore, cptr<Real> poly, const double* bermv, int ring,
                                              int priceRows, int slots, Real* col, double* state, int stride,
                                              Emit&& emit) {
  constexpr bool kRegs = groupUnrolled(G);
  const CCR_CONST GlobalDev& g = *gp;
  const CCR_CONST SetDev& S0 = sets[s0];
  Real* xs = col;
  const Real sgn = (g.antithetic && (pe & 1)) ? Real(-1) : Real(1);
  // The book's shared steps (every set of a book has the same step rows).
  cptr<Real> st0 = steps + static_cast<int64_t>(S0.date_off) * kStepStride;
  Market<Real, Sh> mk;
  marketInit<Real, Sh, false>(mk, g, zbase, S0.num_dates);
  SetPath<Real> regs[kRegs ? G : 1];
#pragma unroll
  for (int k = 0; k < (kRegs ? G : 1); ++k) setPathInit(regs[k]);
  if constexpr (!kRegs) {
    for (int k = 0; k < ns; ++k) {
      state[(kSetStateRows * k) * stride] = 0.0;
      state[(kSetStateRows * k + 1) * stride] = 0.0;
    }
  }
  for (int j = 0; j < S0.num_dates; ++j) {
    if (j > 0) marketStep<Real, Sh, false>(mk, g, j, S0.num_dates, st0, zbase, sgn);
    const CCR_CONST DateRec& r0 = dates[S0.date_off + j];
    storeFixings(mk, r0.fixstore_beg, r0.fixstore_end, fstore + 2 * S0.fixstore_off, xs, stride);
    // Set k of the group on date j: value, collateral and (reporting dates) statistics.
    auto setStep = [&](int k, SetPath<Real>& st) {
      const CCR_CONST SetDev& S = sets[s0 + k];
      const CCR_CONST DateRec& rec = dates[S.date_off + j];
      if (!(rec.flags & kValued) && !S.has_csa) return;
      cptr<Real> dr = dreal + (static_cast<int64_t>(S.date_off) + j) * kDateRealStride;
      Real V = 0;
      if (rec.flags & kValued) {
        V = valueRecord<Real, Sh, false>(mk, g, rec, dr, poly + S.poly_off, bonds + 2 * static_cast<int64_t>(S.bond_off),
                                         fixed + 4 * static_cast<int64_t>(S.fixed_off), nullptr, xs, stride);
        // Bermudan swaptions: the set's row of this valued date from the valuation kernel.
        if (bermv && S.berm_end > S.berm_beg)
          V += static_cast<Real>(bermv[(static_cast<int64_t>(S.cd_base) + rec.vidx) * g.num_paths + pe]);
      }
      Real* ringV = col + (slots + k * (ring + priceRows)) * stride;
      collateralAndReport(mk, S, rec, dr, j, V, st, ringV, ringV + ring * stride, stride,
                          [&](int rr, const double* st9, double vp) { emit(k, rr, st9, vp); });
    };
    if constexpr (kRegs) {
#pragma unroll
      for (int k = 0; k < G; ++k)
        if (k < ns) setStep(k, regs[k]);
    } else {
      for (int k = 0; k < ns; ++k) {
        SetPath<Real> st;
        st.units = static_cast<Real>(state[(kSetStateRows * k) * stride]);
        st.prevV = state[(kSetStateRows * k + 1) * stride];
        setStep(k, st);
        state[(kSetStateRows * k) * stride] = static_cast<double>(st.units);
        state[(kSetStateRows * k + 1) * stride] = st.prevV;
      }
    }
    // A book's sets report on the same dates (the book's reporting grid).
    if (r0.flags & kReporting) marketReported(mk);
  }
}

// Convenience: the tables of set S for path pe, from a PathArgs bundle. Per-trade values take the
// LDS rows after the fixing states.
template <typename Real, typename Sh, bool kOptions = true, bool kTrades = false, typename Emit, typename OnValue,
          typename OnTrades>
__host__ __device__ inline void simulatePath(const PathArgs<Real>& a, cptr<SetDev> sp, int pe, Real* col, int stride,
                                             Emit&& emit, OnValue&& onValue, OnTrades&& onTrades) {
  const CCR_CONST SetDev& S = *sp;
  cptr<GlobalDev> g = toConst(a.g);
  const int64_t q = g->antithetic ? (pe >> 1) : pe;
  const double* bermv = S.berm_end > S.berm_beg ? a.bermv : nullptr;
  simulatePath<Real, Sh, kOptions, kTrades>(
      g, sp, pe, a.stream + q * S.stream_stride, toConst(a.steps) + static_cast<int64_t>(S.date_off) * kStepStride,
      toConst(a.dates) + S.date_off, toConst(a.date_reals) + static_cast<int64_t>(S.date_off) * kDateRealStride,
      toConst(a.bonds) + 2 * static_cast<int64_t>(S.bond_off), toConst(a.fixed) + 4 * static_cast<int64_t>(S.fixed_off),
      toConst(a.fixstore) + 2 * S.fixstore_off, toConst(a.poly) + S.poly_off,
      toConst(a.opts) + kOptStride * static_cast<int64_t>(S.opt_off), bermv, a.ring, a.price_rows, col, stride, emit,
      onValue, toConst(a.tdates) + S.tdate_off, toConst(a.tdreals) + static_cast<int64_t>(S.tdate_off) * kDateRealStride,
      col + (a.ring + a.price_rows + a.slots) * stride, onTrades);
}
template <typename Real, typename Sh, bool kOptions = true, typename Emit, typename OnValue>
__host__ __device__ inline void simulatePath(const PathArgs<Real>& a, cptr<SetDev> sp, int pe, Real* col, int stride,
                                             Emit&& emit, OnValue&& onValue) {
  simulatePath<Real, Sh, kOptions, false>(a, sp, pe, col, stride, emit, onValue,
                                          [](int, Real, Real, const Real*, int) {});
}
template <typename Real, typename Sh, bool kOptions = true, typename Emit>
__host__ __device__ inline void simulatePath(const PathArgs<Real>& a, cptr<SetDev> sp, int pe, Real* col, int stride,
                                             Emit&& emit) {
  simulatePath<Real, Sh, kOptions>(a, sp, pe, col, stride, emit, [](int, Real, Real) {});
}

// The AMC states pass for path pe of set S: the domestic state x and the deflator D of every
// date, written to a.states ([state_off + j P + pe], D after a.state_len elements) when write
// is set. Only the domestic rate is simulated: its driver is the first correlated normal (the
// Cholesky factor's first row has one entry) plus its integral normal, and x and D depend on
// nothing else; the arithmetic is that of simulatePath's domestic rate, so the values are the
// ones the exposure pass sees.
template <typename Real>
__host__ __device__ inline void simulateStates(const PathArgs<Real>& a, cptr<SetDev> sp, int pe, bool write) {
  constexpr bool kDouble = std::is_same<Real, double>::value;
  const CCR_CONST SetDev& S = *sp;
  const CCR_CONST GlobalDev& g = *toConst(a.g);
  if (S.state_off < 0 || !a.states) return;
  const int64_t paths = g.num_paths;
  const int64_t q = g.antithetic ? (pe >> 1) : pe;
  // The two normals of a step: from the gathered domestic normals ([step][2][path block]) when
  // the plan holds them, else from the stream itself.
  const int64_t Q = g.antithetic ? (paths + 1) / 2 : paths;
  const bool gathered = a.dom && S.dom_off >= 0;
  const Real* zb = gathered ? a.dom + S.dom_off + q : a.stream + q * S.stream_stride;
  const int64_t zStep = gathered ? 2 * Q : g.nz, zSecond = gathered ? Q : g.ncorr;
  cptr<Real> steps = toConst(a.steps) + static_cast<int64_t>(S.date_off) * kStepStride;
  const Real sgn = (g.antithetic && (pe & 1)) ? Real(-1) : Real(1);
  const Real c00 = static_cast<Real>(g.chol[0]);
  double* xs = a.states + S.state_off + pe;
  double* ds = xs + a.state_len;
  Real x = 0, D = 1, lnD = 0, lnDc = 0;
  for (int j = 0; j < S.num_dates; ++j) {
    if (j > 0) {
      cptr<Real> o = steps + static_cast<int64_t>(j) * kStepStride + kStepRate;
      const Real* z = zb + static_cast<int64_t>(j - 1) * zStep;
      const Real z1 = sgn * z[0], z2 = sgn * z[zSecond];
      Real w = 0;
      w += c00 * z1;
      const Real x0 = x;
      x = x0 * o[0] + o[1] * w + o[5];
      const Real ix = x0 * o[4] + o[2] * w + o[3] * z2 + o[6];
      const Real ir = ix + o[7];
      if constexpr (kDouble) {
        D *= exp(-ir);
      } else {
        const Real yk = ir - lnDc;
        const Real tk = lnD + yk;
        lnDc = (tk - lnD) - yk;
        lnD = tk;
        D = expr(-lnD);
      }
    }
    if (write) {
      xs[static_cast<int64_t>(j) * paths] = static_cast<double>(x);
      ds[static_cast<int64_t>(j) * paths] = static_cast<double>(D);
    }
  }
}

}  // namespace ccrhip
