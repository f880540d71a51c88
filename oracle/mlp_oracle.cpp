// mlp_oracle.cpp -- TEST INFRASTRUCTURE ONLY: the FP64 CPU checker of the MLP surrogate. Never
// linked into the product.
//
// Restates the reference's surrogate training (src/ai/surrogate.cpp):
//   Standardiser::fit / apply / invert              (:64-90)
//   Mulberry32 and the Box-Muller N(0, 1/fan_in) init (:17-27, :153-162)
//   forward: activation on hidden layers, linear output (:138-152)
//   fitMlp: MSE, gradient summed over fixed 16-sample chunks in chunk order, full-batch Adam
//           (beta 0.9 / 0.999, eps 1e-8, bias-corrected, weight decay on weights only) (:143-245)
// The reference network has two tanh hidden layers; the restatement takes any number of hidden
// layers and SiLU as well (extension: pinned by the reference architecture when restricted to
// it, and by central finite differences of the loss for SiLU and deeper networks).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/ccr_hip.h"

namespace {

struct Layout {
  std::vector<int> sizes;           // inputs, hidden x L, outputs
  std::vector<int64_t> offW, offB;  // per weight layer
  int64_t count = 0;
};

Layout layoutOf(const ccr_mlp_desc& d) {
  Layout L;
  L.sizes.push_back(d.inputs);
  for (int l = 0; l < d.hidden_layers; ++l) L.sizes.push_back(d.hidden);
  L.sizes.push_back(d.outputs);
  for (std::size_t l = 0; l + 1 < L.sizes.size(); ++l) {
    L.offW.push_back(L.count);
    L.count += static_cast<int64_t>(L.sizes[l + 1]) * L.sizes[l];
    L.offB.push_back(L.count);
    L.count += L.sizes[l + 1];
  }
  return L;
}

// mulberry32 (surrogate.cpp:17-27).
struct Mulberry32 {
  uint32_t a;
  double operator()() {
    a += 0x6D2B79F5u;
    uint32_t t = a;
    t = (t ^ (t >> 15)) * (t | 1u);
    t ^= t + (t ^ (t >> 7)) * (t | 61u);
    return static_cast<double>(t ^ (t >> 14)) / 4294967296.0;
  }
};

double act(int kind, double s) { return kind == CCR_ACT_TANH ? std::tanh(s) : s / (1.0 + std::exp(-s)); }
// Derivative from the pre-activation s and the activation a.
double actPrime(int kind, double s, double a) {
  if (kind == CCR_ACT_TANH) return 1.0 - a * a;
  const double sg = 1.0 / (1.0 + std::exp(-s));
  return sg * (1.0 + s * (1.0 - sg));
}

struct Acts {
  std::vector<std::vector<double>> a, z;  // activations (a[0] = input) and pre-activations
};

Acts forward(const ccr_mlp_desc& d, const Layout& L, const double* p, const double* x) {
  const std::size_t nl = L.sizes.size() - 1;
  Acts r;
  r.a.resize(nl + 1);
  r.z.resize(nl + 1);
  r.a[0].assign(x, x + d.inputs);
  for (std::size_t l = 0; l < nl; ++l) {
    const int fo = L.sizes[l + 1], fi = L.sizes[l];
    r.a[l + 1].resize(static_cast<std::size_t>(fo));
    r.z[l + 1].resize(static_cast<std::size_t>(fo));
    for (int i = 0; i < fo; ++i) {
      double s = p[L.offB[l] + i];
      for (int j = 0; j < fi; ++j) s += p[L.offW[l] + static_cast<int64_t>(i) * fi + j] * r.a[l][static_cast<std::size_t>(j)];
      r.z[l + 1][static_cast<std::size_t>(i)] = s;
      r.a[l + 1][static_cast<std::size_t>(i)] = l + 1 < nl ? act(d.activation, s) : s;
    }
  }
  return r;
}

// Gradient of the MSE over standardised samples (fitMlp's loop body, surrogate.cpp:185-213):
// sums over fixed chunks of 16 samples, then the chunk sums in chunk order. Returns the loss.
double gradient(const ccr_mlp_desc& d, const Layout& L, const double* p, const double* xs, const double* ys, int64_t n,
                double* g) {
  constexpr int64_t kChunk = 16;
  const std::size_t nl = L.sizes.size() - 1;
  const int nOut = d.outputs;
  std::vector<double> gr(static_cast<std::size_t>(L.count));
  std::fill(g, g + L.count, 0.0);
  double loss = 0.0;
  for (int64_t c0 = 0; c0 < n; c0 += kChunk) {
    std::fill(gr.begin(), gr.end(), 0.0);
    for (int64_t s = c0; s < std::min(n, c0 + kChunk); ++s) {
      const Acts a = forward(d, L, p, xs + s * d.inputs);
      std::vector<double> delta(static_cast<std::size_t>(nOut));
      for (int k = 0; k < nOut; ++k) {
        const double e = a.a[nl][static_cast<std::size_t>(k)] - ys[s * nOut + k];
        delta[static_cast<std::size_t>(k)] = 2.0 * e / static_cast<double>(n * nOut);
        loss += e * e / static_cast<double>(n * nOut);
      }
      for (std::size_t l = nl; l-- > 0;) {
        const auto& prev = a.a[l];
        const std::size_t nPrev = prev.size();
        for (std::size_t i = 0; i < delta.size(); ++i) {
          gr[static_cast<std::size_t>(L.offB[l]) + i] += delta[i];
          for (std::size_t j = 0; j < nPrev; ++j)
            gr[static_cast<std::size_t>(L.offW[l]) + i * nPrev + j] += delta[i] * prev[j];
        }
        if (l > 0) {
          std::vector<double> next(nPrev, 0.0);
          for (std::size_t j = 0; j < nPrev; ++j) {
            double t = 0.0;
            for (std::size_t i = 0; i < delta.size(); ++i) t += p[static_cast<std::size_t>(L.offW[l]) + i * nPrev + j] * delta[i];
            next[j] = t * actPrime(d.activation, a.z[l][j], prev[j]);
          }
          delta = std::move(next);
        }
      }
    }
    for (int64_t q = 0; q < L.count; ++q) g[q] += gr[static_cast<std::size_t>(q)];
  }
  return loss;
}

}  // namespace

extern "C" {

int64_t ccr_oracle_mlp_param_count(const ccr_mlp_desc* d) { return layoutOf(*d).count; }

// Standardiser::fit (surrogate.cpp:64-78): mean and population sd per column, sd 0 -> 1.
void ccr_oracle_standardiser_fit(const double* rows, int64_t n, int32_t dim, double* mean, double* sd) {
  const double nn = static_cast<double>(n);
  for (int j = 0; j < dim; ++j) mean[j] = sd[j] = 0.0;
  for (int64_t r = 0; r < n; ++r)
    for (int j = 0; j < dim; ++j) mean[j] += rows[r * dim + j] / nn;
  for (int64_t r = 0; r < n; ++r)
    for (int j = 0; j < dim; ++j) sd[j] += (rows[r * dim + j] - mean[j]) * (rows[r * dim + j] - mean[j]) / nn;
  for (int j = 0; j < dim; ++j) sd[j] = std::sqrt(sd[j]) > 0.0 ? std::sqrt(sd[j]) : 1.0;
}

// fitMlp's initialisation (surrogate.cpp:153-162): W ~ N(0, 1 / fan_in) row by row from
// Mulberry32 + Box-Muller (the log uniform drawn first), biases 0.
void ccr_oracle_mlp_init(const ccr_mlp_desc* d, uint32_t seed, double* params) {
  const Layout L = layoutOf(*d);
  Mulberry32 rnd{seed};
  auto gauss = [&] {
    const double u1 = rnd();
    const double u2 = rnd();
    return std::sqrt(-2.0 * std::log(u1 + 1e-12)) * std::cos(2.0 * M_PI * u2);
  };
  for (std::size_t l = 0; l + 1 < L.sizes.size(); ++l) {
    const int fo = L.sizes[l + 1], fi = L.sizes[l];
    for (int i = 0; i < fo; ++i)
      for (int j = 0; j < fi; ++j)
        params[L.offW[l] + static_cast<int64_t>(i) * fi + j] = gauss() * std::sqrt(1.0 / static_cast<double>(fi));
    for (int i = 0; i < fo; ++i) params[L.offB[l] + i] = 0.0;
  }
}

// Loss and gradient at `params` over standardised samples.
double ccr_oracle_mlp_gradient(const ccr_mlp_desc* d, const double* params, const double* xs, const double* ys,
                               int64_t n, double* grad) {
  const Layout L = layoutOf(*d);
  return gradient(*d, L, params, xs, ys, n, grad);
}

// fitMlp (surrogate.cpp:143-245) on raw rows: standardisers (unless warm), init (unless warm),
// `epochs` full-batch Adam steps. stdz: x_mean, x_sd, y_mean, y_sd.
void ccr_oracle_fit_mlp(const ccr_mlp_desc* d, const double* X, const double* Y, int64_t n,
                        const ccr_mlp_train_options* o, double* params, double* stdz) {
  const Layout L = layoutOf(*d);
  double* xm = stdz;
  double* xsd = stdz + d->inputs;
  double* ym = stdz + 2 * d->inputs;
  double* ysd = stdz + 2 * d->inputs + d->outputs;
  if (!o->warm) {
    ccr_oracle_standardiser_fit(X, n, d->inputs, xm, xsd);
    ccr_oracle_standardiser_fit(Y, n, d->outputs, ym, ysd);
    ccr_oracle_mlp_init(d, o->seed, params);
  }
  std::vector<double> xs(static_cast<std::size_t>(n * d->inputs)), ys(static_cast<std::size_t>(n * d->outputs));
  for (int64_t r = 0; r < n; ++r) {
    for (int j = 0; j < d->inputs; ++j) xs[static_cast<std::size_t>(r * d->inputs + j)] = (X[r * d->inputs + j] - xm[j]) / xsd[j];
    for (int k = 0; k < d->outputs; ++k) ys[static_cast<std::size_t>(r * d->outputs + k)] = (Y[r * d->outputs + k] - ym[k]) / ysd[k];
  }
  std::vector<double> g(static_cast<std::size_t>(L.count)), mom(g.size(), 0.0), vel(g.size(), 0.0);
  const double b1 = 0.9, b2 = 0.999;
  for (int ep = 1; ep <= o->epochs; ++ep) {
    gradient(*d, L, params, xs.data(), ys.data(), n, g.data());
    const double c1 = 1.0 - std::pow(b1, static_cast<double>(ep)), c2 = 1.0 - std::pow(b2, static_cast<double>(ep));
    auto adam = [&](int64_t q, bool decay) {
      double& w = params[q];
      const double gq = g[static_cast<std::size_t>(q)] + (decay ? o->weight_decay * w : 0.0);
      mom[static_cast<std::size_t>(q)] = b1 * mom[static_cast<std::size_t>(q)] + (1.0 - b1) * gq;
      vel[static_cast<std::size_t>(q)] = b2 * vel[static_cast<std::size_t>(q)] + (1.0 - b2) * gq * gq;
      w -= o->learning_rate * (mom[static_cast<std::size_t>(q)] / c1) / (std::sqrt(vel[static_cast<std::size_t>(q)] / c2) + 1e-8);
    };
    for (std::size_t l = 0; l + 1 < L.sizes.size(); ++l) {
      const int fo = L.sizes[l + 1], fi = L.sizes[l];
      for (int i = 0; i < fo; ++i) {
        for (int j = 0; j < fi; ++j) adam(L.offW[l] + static_cast<int64_t>(i) * fi + j, true);
        adam(L.offB[l] + i, false);
      }
    }
  }
}

// Mlp::predict (surrogate.cpp:157): invert(forward(apply(x))).
void ccr_oracle_mlp_predict(const ccr_mlp_desc* d, const double* params, const double* stdz, const double* X, int64_t n,
                            double* Y) {
  const Layout L = layoutOf(*d);
  const double* xm = stdz;
  const double* xsd = stdz + d->inputs;
  const double* ym = stdz + 2 * d->inputs;
  const double* ysd = stdz + 2 * d->inputs + d->outputs;
  std::vector<double> z(static_cast<std::size_t>(d->inputs));
  for (int64_t r = 0; r < n; ++r) {
    for (int j = 0; j < d->inputs; ++j) z[static_cast<std::size_t>(j)] = (X[r * d->inputs + j] - xm[j]) / xsd[j];
    const Acts a = forward(*d, L, params, z.data());
    for (int k = 0; k < d->outputs; ++k) Y[r * d->outputs + k] = a.a.back()[static_cast<std::size_t>(k)] * ysd[k] + ym[k];
  }
}

}  // extern "C"
