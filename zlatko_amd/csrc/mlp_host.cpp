// mlp_host.cpp -- host side of the MLP surrogate (see mlp_host.hpp).
#include "mlp_host.hpp"

#include <cmath>
#include <string>

#include "model.hpp"

namespace ccrhip {

namespace {

void hipCheck(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(CCR_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  if (e != hipSuccess) fail(CCR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// mulberry32 (surrogate.cpp:17-27).
struct Mulberry32 {
  uint32_t a;
  double next() {
    a += 0x6D2B79F5u;
    uint32_t t = a;
    t = (t ^ (t >> 15)) * (t | 1u);
    t ^= t + (t ^ (t >> 7)) * (t | 61u);
    return static_cast<double>(t ^ (t >> 14)) / 4294967296.0;
  }
};

int padTo(int v, int p) { return (v + p - 1) / p * p; }

}  // namespace

MlpShape mlpShape(const ccr_mlp_desc& d) {
  if (d.inputs <= 0 || d.outputs <= 0 || d.hidden <= 0 || d.hidden_layers < 0)
    invalid("Mlp: sizes must be positive");
  if (d.hidden_layers + 1 > kMlpMaxLayers) fail(CCR_ERR_UNSUPPORTED, "Mlp: at most 5 hidden layers");
  if (d.inputs > kMlpMaxWidth || d.outputs > kMlpMaxWidth || d.hidden > kMlpMaxWidth)
    fail(CCR_ERR_UNSUPPORTED, "Mlp: widths up to 128");
  if (d.activation != CCR_ACT_TANH && d.activation != CCR_ACT_SILU) invalid("Mlp: unknown activation");
  MlpShape sh;
  sh.layers = d.hidden_layers + 1;
  sh.act = d.activation;
  sh.size[0] = d.inputs;
  for (int l = 1; l <= d.hidden_layers; ++l) sh.size[l] = d.hidden;
  sh.size[sh.layers] = d.outputs;
  for (int l = 0; l <= sh.layers; ++l) sh.pad[l] = padTo(sh.size[l], kMlpPad);
  for (int l = 0; l < sh.layers; ++l) {
    sh.offW[l] = sh.count;
    sh.count += static_cast<int64_t>(sh.size[l + 1]) * sh.size[l];
    sh.offB[l] = sh.count;
    sh.count += sh.size[l + 1];
  }
  return sh;
}

void standardiserFit(const double* rows, int64_t n, int dim, double* mean, double* sd) {
  const double nn = static_cast<double>(n);
  for (int j = 0; j < dim; ++j) mean[j] = sd[j] = 0.0;
  for (int64_t r = 0; r < n; ++r)
    for (int j = 0; j < dim; ++j) mean[j] += rows[r * dim + j] / nn;
  for (int64_t r = 0; r < n; ++r)
    for (int j = 0; j < dim; ++j) {
      const double c = rows[r * dim + j] - mean[j];
      sd[j] += c * c / nn;
    }
  for (int j = 0; j < dim; ++j) sd[j] = std::sqrt(sd[j]) > 0.0 ? std::sqrt(sd[j]) : 1.0;
}

void mlpInit(const MlpShape& sh, uint32_t seed, double* params) {
  Mulberry32 rnd{seed};
  for (int l = 0; l < sh.layers; ++l) {
    const int fo = sh.size[l + 1], fi = sh.size[l];
    const double scale = std::sqrt(1.0 / static_cast<double>(fi));
    for (int64_t q = 0; q < static_cast<int64_t>(fo) * fi; ++q) {
      const double u1 = rnd.next();
      const double u2 = rnd.next();
      params[sh.offW[l] + q] = std::sqrt(-2.0 * std::log(u1 + 1e-12)) * std::cos(2.0 * M_PI * u2) * scale;
    }
    for (int i = 0; i < fo; ++i) params[sh.offB[l] + i] = 0.0;
  }
}

void* MlpRun::alloc(size_t bytes) {
  void* p = nullptr;
  hipCheck(hipMalloc(&p, bytes < 256 ? 256 : bytes), "hipMalloc");
  bufs_.push_back(p);
  hipCheck(hipMemsetAsync(p, 0, bytes < 256 ? 256 : bytes, stream_), "hipMemsetAsync");
  return p;
}

MlpRun::MlpRun(const ccr_mlp_desc& d, int precision, int64_t n, hipStream_t stream) : stream_(stream) {
  if (n <= 0) invalid("Mlp: no samples");
  if (precision != CCR_FP32 && precision != CCR_BF16) invalid("Mlp: precision must be CCR_FP32 or CCR_BF16");
  m_.sh = mlpShape(d);
  const MlpShape& sh = m_.sh;
  const int L = sh.layers;
  if (precision == CCR_BF16) {
    if (d.hidden_layers > 0 && d.hidden != 128) fail(CCR_ERR_UNSUPPORTED, "Mlp BF16: hidden width must be 128");
    if (d.inputs > 32) fail(CCR_ERR_UNSUPPORTED, "Mlp BF16: at most 32 inputs");
    if (d.hidden_layers == 0 && d.outputs > 32) fail(CCR_ERR_UNSUPPORTED, "Mlp BF16: at most 32 outputs without hidden layers");
  }
  m_.precision = precision;
  m_.n = n;
  m_.n_pad = precision == CCR_BF16 ? (n + kMlpRows - 1) / kMlpRows * kMlpRows : n;
  m_.ld = m_.n_pad;
  const int64_t np = m_.n_pad;
  m_.nchunks = (np + kMlpChunk - 1) / kMlpChunk;
  try {
    m_.params = static_cast<float*>(alloc(sh.count * sizeof(float)));
    m_.grad = static_cast<float*>(alloc(sh.count * sizeof(float)));
    m_.mom = static_cast<float*>(alloc(sh.count * sizeof(float)));
    m_.vel = static_cast<float*>(alloc(sh.count * sizeof(float)));
    m_.partial = static_cast<float*>(alloc(static_cast<size_t>(m_.nchunks * sh.count) * sizeof(float)));
    m_.ys = static_cast<float*>(alloc(static_cast<size_t>(n * sh.size[L]) * sizeof(float)));
    m_.loss_blocks = static_cast<int>((n * sh.size[L] + 255) / 256);
    m_.loss_part = static_cast<float*>(alloc(static_cast<size_t>(m_.loss_blocks) * sizeof(float)));
    if (precision == CCR_FP32) {
      m_.xs = static_cast<float*>(alloc(static_cast<size_t>(n * sh.size[0]) * sizeof(float)));
      m_.a_sm[0] = m_.xs;
      for (int l = 1; l <= L; ++l) {
        const size_t bytes = static_cast<size_t>(n * sh.size[l]) * sizeof(float);
        m_.a_sm[l] = alloc(bytes);
        m_.z[l] = static_cast<float*>(alloc(bytes));
        m_.d_sm[l] = alloc(bytes);
      }
      std::vector<float*> tab(static_cast<size_t>(2 * (L + 1)), nullptr);
      for (int l = 0; l <= L; ++l) {
        tab[static_cast<size_t>(l)] = static_cast<float*>(m_.a_sm[l]);
        tab[static_cast<size_t>(L + 1 + l)] = static_cast<float*>(m_.d_sm[l]);
      }
      m_.tables = static_cast<float**>(alloc(tab.size() * sizeof(float*)));
      hipCheck(hipMemcpyAsync(m_.tables, tab.data(), tab.size() * sizeof(float*), hipMemcpyHostToDevice, stream_),
               "hipMemcpyAsync");
    } else {
      m_.xs = static_cast<float*>(alloc(static_cast<size_t>(n * sh.size[0]) * sizeof(float)));
      for (int l = 0; l <= L; ++l) {
        const size_t elems = static_cast<size_t>(np * sh.pad[l]);
        if (l < L) {
          m_.a_sm[l] = alloc(elems * 2);
          m_.a_fm[l] = alloc(elems * 2);
        }
        if (l > 0) {
          m_.z[l] = static_cast<float*>(alloc(elems * sizeof(float)));
          m_.d_sm[l] = alloc(elems * 2);
          m_.d_fm[l] = alloc(elems * 2);
        }
      }
      for (int l = 0; l < L; ++l) {
        const size_t w = static_cast<size_t>(sh.pad[l]) * sh.pad[l + 1];
        m_.w16[l] = alloc(w * 2);
        m_.wt16[l] = alloc(w * 2);
        m_.bpad[l] = static_cast<float*>(alloc(static_cast<size_t>(sh.pad[l + 1]) * sizeof(float)));
      }
    }
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  } catch (...) {
    for (void* p : bufs_) (void)hipFree(p);
    bufs_.clear();
    throw;
  }
}

MlpRun::~MlpRun() {
  (void)hipStreamSynchronize(stream_);
  for (void* p : bufs_) (void)hipFree(p);
}

void MlpRun::setParams(const double* params) {
  std::vector<float> f(static_cast<size_t>(m_.sh.count));
  for (size_t i = 0; i < f.size(); ++i) f[i] = static_cast<float>(params[i]);
  hipCheck(hipMemcpyAsync(m_.params, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  hipCheck(mlpRefreshBf16(m_, stream_), "MLP weight copies");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void MlpRun::getParams(double* params) const {
  std::vector<float> f(static_cast<size_t>(m_.sh.count));
  hipCheck(hipMemcpyAsync(f.data(), m_.params, f.size() * sizeof(float), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  for (size_t i = 0; i < f.size(); ++i) params[i] = f[i];
}

void MlpRun::setInputs(const double* xs) {
  std::vector<float> f(static_cast<size_t>(m_.n * m_.sh.size[0]));
  for (size_t i = 0; i < f.size(); ++i) f[i] = static_cast<float>(xs[i]);
  hipCheck(hipMemcpyAsync(m_.xs, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  hipCheck(mlpLoadInputs(m_, m_.xs, stream_), "MLP inputs");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void MlpRun::setTargets(const double* ys) {
  std::vector<float> f(static_cast<size_t>(m_.n * m_.sh.size[m_.sh.layers]));
  for (size_t i = 0; i < f.size(); ++i) f[i] = static_cast<float>(ys[i]);
  hipCheck(hipMemcpyAsync(m_.ys, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void MlpRun::forward() { hipCheck(mlpForward(m_, stream_), "MLP forward"); }

double MlpRun::loss() {
  forward();
  hipCheck(mlpLossGrad(m_, stream_), "MLP loss");
  std::vector<float> lp(static_cast<size_t>(m_.loss_blocks));
  hipCheck(hipMemcpyAsync(lp.data(), m_.loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  double s = 0.0;
  for (float v : lp) s += v;
  return s;
}

double MlpRun::lossAndGradient(double* grad) {
  const double l = loss();
  hipCheck(mlpBackward(m_, stream_), "MLP backward");
  std::vector<float> g(static_cast<size_t>(m_.sh.count));
  hipCheck(hipMemcpyAsync(g.data(), m_.grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  for (size_t i = 0; i < g.size(); ++i) grad[i] = g[i];
  return l;
}

// Samples [b0, b0 + count) of m as a device of their own (BF16 path; b0 a multiple of kMlpRows).
MlpDevice batchView(const MlpDevice& m, int64_t b0, int64_t count) {
  MlpDevice v = m;
  v.n = count;
  v.n_pad = (count + kMlpRows - 1) / kMlpRows * kMlpRows;
  v.nchunks = (v.n_pad + kMlpChunk - 1) / kMlpChunk;
  v.loss_blocks = static_cast<int>((count * m.sh.size[m.sh.layers] + 255) / 256);
  for (int l = 0; l <= m.sh.layers; ++l) {
    const int64_t sm = b0 * m.sh.pad[l];
    if (v.a_sm[l]) v.a_sm[l] = static_cast<__bf16*>(v.a_sm[l]) + sm;
    if (v.a_fm[l]) v.a_fm[l] = static_cast<__bf16*>(v.a_fm[l]) + b0;
    if (v.z[l]) v.z[l] += sm;
    if (v.d_sm[l]) v.d_sm[l] = static_cast<__bf16*>(v.d_sm[l]) + sm;
    if (v.d_fm[l]) v.d_fm[l] = static_cast<__bf16*>(v.d_fm[l]) + b0;
  }
  v.ys += b0 * m.sh.size[m.sh.layers];
  return v;
}

void MlpRun::train(int epochs, double lr, double wd, int64_t batch) {
  if (batch > 0 && m_.precision != CCR_BF16) invalid("Mlp: mini-batches need the BF16 path");
  if (batch > 0 && batch % kMlpRows != 0) invalid("Mlp: the batch size must be a multiple of 64");
  const bool full = batch <= 0 || batch >= m_.n;
  auto step = [&](const MlpDevice& d) {
    ++epoch_;
    hipCheck(mlpForward(d, stream_), "MLP forward");
    hipCheck(mlpLossGrad(d, stream_), "MLP loss");
    hipCheck(mlpBackward(d, stream_), "MLP backward");
    const double c1 = 1.0 - std::pow(0.9, static_cast<double>(epoch_));
    const double c2 = 1.0 - std::pow(0.999, static_cast<double>(epoch_));
    hipCheck(mlpAdam(d, static_cast<float>(lr), static_cast<float>(wd), static_cast<float>(c1), static_cast<float>(c2),
                     stream_),
             "MLP Adam");
  };
  for (int e = 0; e < epochs; ++e) {
    if (full) {
      step(m_);
      continue;
    }
    for (int64_t b0 = 0; b0 < m_.n; b0 += batch) step(batchView(m_, b0, std::min(batch, m_.n - b0)));
  }
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void MlpRun::outputs(double* ys) const {
  const int L = m_.sh.layers, nout = m_.sh.size[L];
  if (m_.precision == CCR_FP32) {
    std::vector<float> f(static_cast<size_t>(m_.n * nout));
    hipCheck(hipMemcpyAsync(f.data(), m_.a_sm[L], f.size() * sizeof(float), hipMemcpyDeviceToHost, stream_),
             "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (size_t i = 0; i < f.size(); ++i) ys[i] = f[i];
    return;
  }
  const int po = m_.sh.pad[L];
  std::vector<float> f(static_cast<size_t>(m_.n * po));
  hipCheck(hipMemcpyAsync(f.data(), m_.z[L], f.size() * sizeof(float), hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  for (int64_t s = 0; s < m_.n; ++s)
    for (int k = 0; k < nout; ++k) ys[s * nout + k] = f[static_cast<size_t>(s * po + k)];
}

}  // namespace ccrhip
