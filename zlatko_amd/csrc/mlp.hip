// mlp.hip -- gfx950 kernels of the MLP surrogate (forward, backward, Adam).
//
// BF16 path (hidden = 128): every product is v_mfma_f32_16x16x32_bf16 on 16-row tiles.
//   Lane layouts (checked on gfx950): A[m][k] -> lane (m + 16 (k / 8) % 64), element k % 8;
//   B[k][n] -> lane (n + 16 (k / 8)), element k % 8; D[m][n] -> lane (n + 16 (m / 4)), element m % 4.
//   forward       Z_{l+1} = A_l W_l^T + b_l      A = sample-major activations, B = W_l rows
//   backward-data D_l = (D_{l+1} W_l) . act'(Z_l)  A = sample-major deltas,      B = W_l^T rows
//   weight grad   dW_l = D_{l+1}^T A_l           A = feature-major deltas,     B = feature-major activations
//   Each weight-gradient partial covers kMlpChunk samples; partials are summed in chunk order.
// FP32 path (any width <= 128): one thread per output element, sums in the reference's order
// (bias first, then inputs ascending; backward over outputs ascending).
// Adam follows fitMlp (src/ai/surrogate.cpp:223-240) in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "mlp.hpp"
#include "wave.hpp"

namespace ccrhip {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ inline float actF(int kind, float z) {
  return kind == CCR_ACT_TANH ? tanhf(z) : z / (1.0f + __expf(-z));
}
// The activation for the fused BF16 inference, whose outputs are rounded to BF16 (8-bit
// mantissa): SiLU with the hardware reciprocal (1 ulp) instead of an IEEE division.
__device__ inline float actInfer(int kind, float z) {
  return kind == CCR_ACT_TANH ? tanhf(z) : z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
// Derivative of the activation at pre-activation z.
__device__ inline float actPrimeF(int kind, float z) {
  if (kind == CCR_ACT_TANH) {
    const float a = tanhf(z);
    return 1.0f - a * a;
  }
  const float sg = 1.0f / (1.0f + __expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

__device__ inline bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// ---------------------------------------------------------------- FP32 path

__global__ void f32_fwd_kernel(const float* ain, int fi, const float* W, const float* b, int fo, float* z, float* aout,
                               int64_t n, int act, int last) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * fo) return;
  const int64_t s = idx / fo;
  const int o = static_cast<int>(idx % fo);
  float v = b[o];
  const float* wr = W + static_cast<int64_t>(o) * fi;
  const float* ar = ain + s * fi;
  for (int j = 0; j < fi; ++j) v += wr[j] * ar[j];
  z[idx] = v;
  aout[idx] = last ? v : actF(act, v);
}

// Output deltas 2 (y - t) / (n out) and per-block squared-error sums.
__global__ void __launch_bounds__(256) f32_loss_kernel(const float* out, const float* ys, float* dout, int64_t n, int nout,
                                                       float* loss_part) {
  __shared__ double lds[4];
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (idx < n * nout) {
    const float e = out[idx] - ys[idx];
    dout[idx] = 2.0f * e / static_cast<float>(n * nout);
    v[0] = static_cast<double>(e) * e;
  }
  blockSum<1>(v, lds, 4);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = static_cast<float>(v[0] / static_cast<double>(n * nout));
}

__global__ void f32_bwd_kernel(const float* dnext, int fo, const float* W, int fi, const float* z, float* d, int64_t n,
                               int act) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * fi) return;
  const int64_t s = idx / fi;
  const int j = static_cast<int>(idx % fi);
  float t = 0.0f;
  for (int o = 0; o < fo; ++o) t += W[static_cast<int64_t>(o) * fi + j] * dnext[s * fo + o];
  d[idx] = t * actPrimeF(act, z[idx]);
}

// Gradient partial of chunk blockIdx.y for parameter q: samples of the chunk in order.
__global__ void f32_wgrad_kernel(MlpShape sh, const float* const* acts, const float* const* deltas, int64_t n,
                                 float* partial) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= sh.count) return;
  int l = 0;
  while (l + 1 < sh.layers && q >= sh.offW[l + 1]) ++l;
  const int fi = sh.size[l], fo = sh.size[l + 1];
  const bool isBias = q >= sh.offB[l];
  const int o = isBias ? static_cast<int>(q - sh.offB[l]) : static_cast<int>((q - sh.offW[l]) / fi);
  const int j = isBias ? 0 : static_cast<int>((q - sh.offW[l]) % fi);
  const float* dl = deltas[l + 1];
  const float* al = acts[l];
  const int64_t s0 = static_cast<int64_t>(blockIdx.y) * kMlpChunk;
  const int64_t s1 = s0 + kMlpChunk < n ? s0 + kMlpChunk : n;
  float g = 0.0f;
  if (isBias) {
    for (int64_t s = s0; s < s1; ++s) g += dl[s * fo + o];
  } else {
    for (int64_t s = s0; s < s1; ++s) g += dl[s * fo + o] * al[s * fi + j];
  }
  partial[static_cast<int64_t>(blockIdx.y) * sh.count + q] = g;
}

__global__ void reduce_kernel(const float* partial, int64_t count, int64_t nchunks, float* grad) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  float g = 0.0f;
  for (int64_t c = 0; c < nchunks; ++c) g += partial[c * count + q];
  grad[q] = g;
}

__global__ void adam_kernel(MlpShape sh, float* params, const float* grad, float* mom, float* vel, float lr, float wd,
                            float c1, float c2) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= sh.count) return;
  int l = 0;
  while (l + 1 < sh.layers && q >= sh.offW[l + 1]) ++l;
  const bool decay = q < sh.offB[l];
  const float b1 = 0.9f, b2 = 0.999f;
  float w = params[q];
  const float gq = grad[q] + (decay ? wd * w : 0.0f);
  const float m = b1 * mom[q] + (1.0f - b1) * gq;
  const float v = b2 * vel[q] + (1.0f - b2) * gq * gq;
  mom[q] = m;
  vel[q] = v;
  w -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
  params[q] = w;
}

// ---------------------------------------------------------------- BF16 path

__global__ void bf16_inputs_kernel(const float* xs, int fin, int pad, int64_t n, int64_t n_pad, __bf16* a_sm,
                                   __bf16* a_fm) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n_pad * pad) return;
  const int64_t s = idx / pad;
  const int j = static_cast<int>(idx % pad);
  const float v = (s < n && j < fin) ? xs[s * fin + j] : 0.0f;
  a_sm[idx] = static_cast<__bf16>(v);
  a_fm[static_cast<int64_t>(j) * n_pad + s] = static_cast<__bf16>(v);
}

// BF16 weight copies (zero-padded) and padded FP32 biases of layer l.
__global__ void bf16_refresh_kernel(const float* params, int64_t offW, int64_t offB, int fi, int fo, int pi, int po,
                                    __bf16* w, __bf16* wt, float* bpad) {
  const int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= pi * po) return;
  const int o = idx / pi, j = idx % pi;
  const float v = (o < fo && j < fi) ? params[offW + static_cast<int64_t>(o) * fi + j] : 0.0f;
  w[idx] = static_cast<__bf16>(v);
  wt[static_cast<int64_t>(j) * po + o] = static_cast<__bf16>(v);
  if (j == 0) bpad[o] = o < fo ? params[offB + o] : 0.0f;
}

// Forward of one layer: 4 wavefronts x 16 samples per workgroup; FI, FO padded widths.
template <int FI, int FO, bool kLast>
__global__ void __launch_bounds__(256) bf16_fwd_kernel(const __bf16* ain, const __bf16* w, const float* b, float* z,
                                                       __bf16* aout_sm, __bf16* aout_fm, int64_t n_pad, int act) {
  const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
  const int64_t s0 = static_cast<int64_t>(blockIdx.x) * kMlpRows + wv * 16;
  const int r = lane & 15, g = lane >> 4;
  bf16x8 afr[FI / 32];
#pragma unroll
  for (int ks = 0; ks < FI / 32; ++ks) afr[ks] = load8(ain + (s0 + r) * FI + ks * 32 + 8 * g);
#pragma unroll
  for (int t = 0; t < FO / 16; ++t) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < FI / 32; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[ks], load8(w + (t * 16 + r) * FI + ks * 32 + 8 * g), acc, 0, 0, 0);
    const int o = t * 16 + r;
    const float bo = b[o];
    bf16x4 col;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t s = s0 + 4 * g + i;
      const float v = acc[i] + bo;
      z[s * FO + o] = v;
      if constexpr (!kLast) {
        const __bf16 a = static_cast<__bf16>(actF(act, v));
        aout_sm[s * FO + o] = a;
        col[i] = a;
      }
    }
    if constexpr (!kLast) *reinterpret_cast<bf16x4*>(aout_fm + static_cast<int64_t>(o) * n_pad + s0 + 4 * g) = col;
  }
}

// Output deltas 2 (y - t) / (n out) in BF16 (sample- and feature-major) and squared-error sums. A
// NaN target is a missing one (a netting set that does not value that date): no error, no delta.
__global__ void __launch_bounds__(256) bf16_loss_kernel(const float* out, int po, const float* ys, int nout, int64_t n,
                                                        int64_t n_pad, __bf16* d_sm, __bf16* d_fm, float* loss_part) {
  __shared__ double lds[4];
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (idx < n * nout) {
    const int64_t s = idx / nout;
    const int k = static_cast<int>(idx % nout);
    const float t = ys[idx];
    const float e = t == t ? out[s * po + k] - t : 0.0f;
    const __bf16 d = static_cast<__bf16>(2.0f * e / static_cast<float>(n * nout));
    d_sm[s * po + k] = d;
    d_fm[static_cast<int64_t>(k) * n_pad + s] = d;
    v[0] = static_cast<double>(e) * e;
  }
  blockSum<1>(v, lds, 4);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = static_cast<float>(v[0] / static_cast<double>(n * nout));
}

// Backward-data of one layer: D_l = (D_{l+1} W_l) . act'(Z_l); FI = width of l, FO = width of l + 1.
template <int FI, int FO>
__global__ void __launch_bounds__(256) bf16_bwd_kernel(const __bf16* dnext, const __bf16* wt, const float* zl,
                                                       __bf16* d_sm, __bf16* d_fm, int64_t n_pad, int act) {
  const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
  const int64_t s0 = static_cast<int64_t>(blockIdx.x) * kMlpRows + wv * 16;
  const int r = lane & 15, g = lane >> 4;
  bf16x8 afr[FO / 32];
#pragma unroll
  for (int ks = 0; ks < FO / 32; ++ks) afr[ks] = load8(dnext + (s0 + r) * FO + ks * 32 + 8 * g);
#pragma unroll
  for (int t = 0; t < FI / 16; ++t) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < FO / 32; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[ks], load8(wt + (t * 16 + r) * FO + ks * 32 + 8 * g), acc, 0, 0, 0);
    const int j = t * 16 + r;
    bf16x4 col;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t s = s0 + 4 * g + i;
      const __bf16 d = static_cast<__bf16>(acc[i] * actPrimeF(act, zl[s * FI + j]));
      d_sm[s * FI + j] = d;
      col[i] = d;
    }
    *reinterpret_cast<bf16x4*>(d_fm + static_cast<int64_t>(j) * n_pad + s0 + 4 * g) = col;
  }
}

// Weight-gradient partial of one 16 x 16 tile of dW_l over one chunk of samples (one wavefront):
// dW[o][j] = sum_s D_{l+1}[o][s] A_l[j][s] from the feature-major copies; tiles of the first column
// also sum the bias gradient. Writes only real (unpadded) entries.
__global__ void __launch_bounds__(64) bf16_wgrad_kernel(const __bf16* dfm, const __bf16* afm, int64_t n_pad, int64_t ld,
                                                        int fi,
                                                        int fo, int pi, int64_t offW, int64_t offB, int64_t count,
                                                        float* partial) {
  const int lane = static_cast<int>(threadIdx.x);
  const int tilesJ = pi / 16;
  const int mt = static_cast<int>(blockIdx.x) / tilesJ, nt = static_cast<int>(blockIdx.x) % tilesJ;
  const int r = lane & 15, g = lane >> 4;
  const int64_t c = blockIdx.y;
  const int64_t k0 = c * kMlpChunk, k1 = k0 + kMlpChunk < n_pad ? k0 + kMlpChunk : n_pad;
  const __bf16* arow = dfm + static_cast<int64_t>(mt * 16 + r) * ld + 8 * g;
  const __bf16* brow = afm + static_cast<int64_t>(nt * 16 + r) * ld + 8 * g;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  for (int64_t k = k0; k < k1; k += 32) {
    const bf16x8 a = load8(arow + k);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, load8(brow + k), acc, 0, 0, 0);
    if (nt == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) bsum += static_cast<float>(a[i]);
    }
  }
  float* out = partial + c * count;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = mt * 16 + 4 * g + i, j = nt * 16 + r;
    if (o < fo && j < fi) out[offW + static_cast<int64_t>(o) * fi + j] = acc[i];
  }
  if (nt == 0) {
    // Lanes r, r + 16, r + 32, r + 48 hold the four sample groups of output row mt * 16 + r.
    const float s1 = __shfl(bsum, r + 16), s2 = __shfl(bsum, r + 32), s3 = __shfl(bsum, r + 48);
    const int o = mt * 16 + r;
    if (g == 0 && o < fo) out[offB + o] = ((bsum + s1) + s2) + s3;
  }
}

// Fused inference (hidden = 128): every layer of kTiles 16-sample tiles per wavefront, the weights
// and biases staged once per workgroup in LDS. Each product is computed transposed, Z^T = W A^T: a
// weight fragment (A operand, 16 outputs x 32 inputs) is read from LDS once and multiplied into
// every sample tile of the wavefront (B operands, 32 inputs x 16 samples), and the accumulators
// start at the bias. A layer's 8 x kTiles accumulator tiles are all computed before any is
// activated, so the MFMA chains and the activations of different tiles interleave. The result
// lands as four consecutive outputs of one sample per lane, so a layer's activations go back to the
// wavefront's LDS tile in one 8-byte store per lane and tile, sample-major, where the next layer
// reads its B operands (16 bytes per lane). Weight rows in LDS are padded by 8 elements so that the
// 16 lanes reading the rows of one fragment hit distinct banks (unless the padding would leave no
// room for four wavefronts: five hidden layers). Inputs x [n][inputs] (raw) are
// standardised on load with (mean, sd). The last layer's outputs (standardised) go to the sink:
// y [n][outputs], or Bermudan value rows (below). Grid-stride over groups of 16 kTiles samples. The
// wavefronts per workgroup are chosen at launch so that the weights and one activation tile per
// wavefront fit the workgroup's LDS.
constexpr int kActStride = 128 + 8;  // BF16 elements per activation row in LDS (padding vs bank conflicts)
constexpr size_t kInferLdsBytes = 160 * 1024;  // LDS of one gfx950 workgroup
struct InferArgs {
  const __bf16* w[kMlpMaxLayers];
  const float* b[kMlpMaxLayers];
  const float* x;
  float* y;
  float mean[32], rsd[32];  // input standardiser: (x - mean) * rsd, rsd = 1 / sd
  int64_t n;
  int layers, inputs, outputs;
  int w0stride, wstride;  // LDS rows of the first layer's weights and of the others' (elements)
  // Value-row sink (y == nullptr): sample s = u * paths + p; output o goes to
  // rows[(outs[o].base + row_map[o * urows + u]) * paths + p] = y_o * outs[o].sd + outs[o].mean where
  // the map is >= 0 (dates output o's netting set values), nowhere otherwise.
  double* rows;
  const int32_t* row_map;
  const InferOut* outs;
  int64_t paths;
  int32_t urows;
};

// LDS elements (BF16) of the staged weights and floats of the staged biases.
__host__ __device__ inline int inferWeightElems(int layers, int po, int w0stride, int wstride) {
  return 128 * w0stride + (layers - 2) * 128 * wstride + po * wstride;
}
__host__ __device__ inline int inferBiasFloats(int layers, int po) { return (layers - 1) * 128 + po; }

// Rows x fi weights (row-major, BF16) into LDS rows of `stride` elements.
__device__ inline void stageRows(__bf16* dst, const __bf16* src, int rows, int fi, int stride) {
  const int chunks = rows * (fi / 8);
  for (int c = static_cast<int>(threadIdx.x); c < chunks; c += static_cast<int>(blockDim.x)) {
    const int r = c / (fi / 8), k = (c % (fi / 8)) * 8;
    *reinterpret_cast<bf16x8*>(dst + r * stride + k) = load8(src + r * fi + k);
  }
}

template <int kAct>
__device__ inline float activate(float z) {
  if constexpr (kAct == CCR_ACT_TANH) return tanhf(z);
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// Activations of a layer's 8 x kTiles tiles into the wavefront's LDS tile.
template <int kTiles, int kAct>
__device__ inline void activateStore(const f32x4 (&acc)[8][kTiles], __bf16* act, int n, int g) {
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int tl = 0; tl < kTiles; ++tl) {
      bf16x4 v;
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = static_cast<__bf16>(activate<kAct>(acc[t][tl][i]));
      *reinterpret_cast<bf16x4*>(act + (16 * tl + n) * kActStride + t * 16 + 4 * g) = v;
    }
}

// Wavefronts per workgroup at most: 12 with one tile (168 VGPRs), 8 with two (256 VGPRs).
constexpr int inferMaxWaves(int tiles) { return tiles == 1 ? 12 : 8; }

template <int kTiles, int kAct>
__global__ void __launch_bounds__(64 * inferMaxWaves(kTiles)) bf16_infer_kernel(InferArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __bf16* wl = reinterpret_cast<__bf16*>(lds_raw);
  const int L = a.layers, po = (a.outputs + 31) / 32 * 32, ntl = po / 16;
  // Weights: first 128 x 32, middle 128 x 128 (L - 2 of them), last po x 128; then the biases.
  const int w0s = a.w0stride, ws = a.wstride;
  stageRows(wl, a.w[0], 128, 32, w0s);
  for (int l = 1; l + 1 < L; ++l) stageRows(wl + 128 * w0s + (l - 1) * 128 * ws, a.w[l], 128, 128, ws);
  __bf16* wlast = wl + 128 * w0s + (L - 2) * 128 * ws;
  stageRows(wlast, a.w[L - 1], po, 128, ws);
  const int welems = inferWeightElems(L, po, w0s, ws);
  float* bl = reinterpret_cast<float*>(wl + welems);
  for (int l = 0; l < L; ++l) {
    const int fo = l + 1 == L ? po : 128;
    for (int i = static_cast<int>(threadIdx.x); i < fo; i += static_cast<int>(blockDim.x)) bl[l * 128 + i] = a.b[l][i];
  }
  __bf16* act = wl + welems + 2 * inferBiasFloats(L, po) + (static_cast<int>(threadIdx.x) >> 6) * (16 * kTiles) * kActStride;
  __syncthreads();
  const int lane = static_cast<int>(threadIdx.x) & 63, n = lane & 15, g = lane >> 4;
  const int waves = static_cast<int>(blockDim.x) >> 6;
  const int64_t groups = (a.n + 16 * kTiles - 1) / (16 * kTiles);
  for (int64_t grp = static_cast<int64_t>(blockIdx.x) * waves + (threadIdx.x >> 6); grp < groups;
       grp += static_cast<int64_t>(gridDim.x) * waves) {
    const int64_t s0 = grp * 16 * kTiles;
    f32x4 acc[8][kTiles];
    bf16x8 bfr[kTiles][4];
    // Layer 0: B operands from the inputs (sample s0 + 16 tile + n, features 8 g .. 8 g + 7).
#pragma unroll
    for (int tl = 0; tl < kTiles; ++tl) {
      const int64_t s = s0 + 16 * tl + n;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int j = 8 * g + i;
        float v = 0.0f;
        if (j < a.inputs && s < a.n) v = (a.x[s * a.inputs + j] - a.mean[j]) * a.rsd[j];
        bfr[tl][0][i] = static_cast<__bf16>(v);
      }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const f32x4 bias = *reinterpret_cast<const f32x4*>(bl + t * 16 + 4 * g);
      const bf16x8 wf = *reinterpret_cast<const bf16x8*>(wl + (t * 16 + n) * w0s + 8 * g);
#pragma unroll
      for (int tl = 0; tl < kTiles; ++tl) acc[t][tl] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, bfr[tl][0], bias, 0, 0, 0);
    }
    activateStore<kTiles, kAct>(acc, act, n, g);
    // Hidden 128 x 128 layers (every lane reads its B operands before any lane of the wavefront
    // overwrites the tile: wavefront-synchronous, LDS operations in order).
    for (int l = 1; l + 1 < L; ++l) {
      const __bf16* w = wl + 128 * w0s + (l - 1) * 128 * ws;
#pragma unroll
      for (int tl = 0; tl < kTiles; ++tl)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          bfr[tl][ks] = *reinterpret_cast<const bf16x8*>(act + (16 * tl + n) * kActStride + ks * 32 + 8 * g);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const f32x4 bias = *reinterpret_cast<const f32x4*>(bl + l * 128 + t * 16 + 4 * g);
#pragma unroll
        for (int tl = 0; tl < kTiles; ++tl) acc[t][tl] = bias;
      }
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const bf16x8 wf = *reinterpret_cast<const bf16x8*>(w + (t * 16 + n) * ws + ks * 32 + 8 * g);
#pragma unroll
          for (int tl = 0; tl < kTiles; ++tl)
            acc[t][tl] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, bfr[tl][ks], acc[t][tl], 0, 0, 0);
        }
      activateStore<kTiles, kAct>(acc, act, n, g);
    }
    // Last layer: po outputs (ntl tiles of 16).
#pragma unroll
    for (int tl = 0; tl < kTiles; ++tl)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        bfr[tl][ks] = *reinterpret_cast<const bf16x8*>(act + (16 * tl + n) * kActStride + ks * 32 + 8 * g);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t >= ntl) break;
      const f32x4 bias = *reinterpret_cast<const f32x4*>(bl + (L - 1) * 128 + t * 16 + 4 * g);
#pragma unroll
      for (int tl = 0; tl < kTiles; ++tl) acc[t][tl] = bias;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 wf = *reinterpret_cast<const bf16x8*>(wlast + (t * 16 + n) * ws + ks * 32 + 8 * g);
#pragma unroll
        for (int tl = 0; tl < kTiles; ++tl)
          acc[t][tl] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, bfr[tl][ks], acc[t][tl], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      if (t >= ntl) break;
      const int o0 = t * 16 + 4 * g;  // this lane's four outputs o0 .. o0 + 3
#pragma unroll
      for (int tl = 0; tl < kTiles; ++tl) {
        const int64_t s = s0 + 16 * tl + n;
        if (s >= a.n) continue;
        if (a.y) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (o0 + i < a.outputs) a.y[s * a.outputs + o0 + i] = acc[t][tl][i];
        } else {
          const int64_t u = s / a.paths, p = s - u * a.paths;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int o = o0 + i;
            if (o >= a.outputs) continue;
            const int32_t r = a.row_map[static_cast<int64_t>(o) * a.urows + u];
            if (r < 0) continue;
            const InferOut oc = a.outs[o];
            a.rows[(oc.base + static_cast<int64_t>(r)) * a.paths + p] = static_cast<double>(acc[t][tl][i]) * oc.sd + oc.mean;
          }
        }
      }
    }
  }
}

// Launch shape of the fused inference: sample tiles per wavefront and wavefronts per workgroup
// (two tiles while at least four wavefronts fit beside the weights, else one), and the LDS bytes.
struct InferShape {
  int tiles = 0, waves = 0, w0stride = 0, wstride = 0;
  size_t lds = 0;
};
// Padded weight rows (conflict-free fragment reads) with two tiles per wavefront while at least
// four wavefronts fit, else unpadded rows, two tiles and four wavefronts, else one tile.
InferShape inferShape(const MlpShape& sh) {
  const int po = sh.pad[sh.layers];
  InferShape r;
  const int strides[2][2] = {{32 + 8, 128 + 8}, {32, 128}};
  for (int tiles : {1, 2})
    for (const auto& st : strides) {
      const size_t fixed = static_cast<size_t>(inferWeightElems(sh.layers, po, st[0], st[1])) * 2 +
                           static_cast<size_t>(inferBiasFloats(sh.layers, po)) * 4;
      const size_t tile = static_cast<size_t>(16 * tiles) * kActStride * 2;
      if (fixed + tile > kInferLdsBytes) continue;
      const int waves = static_cast<int>(std::min<size_t>(inferMaxWaves(tiles), (kInferLdsBytes - fixed) / tile));
      if (tiles == 2 && waves < 4) continue;
      r.tiles = tiles;
      r.waves = waves;
      r.w0stride = st[0];
      r.wstride = st[1];
      r.lds = fixed + static_cast<size_t>(waves) * tile;
      return r;
    }
  return r;
}

template <int kTiles>
void inferLaunch(const InferArgs& a, int act, const InferShape& is, unsigned grid, hipStream_t stream) {
  if (act == CCR_ACT_TANH)
    hipLaunchKernelGGL((bf16_infer_kernel<kTiles, CCR_ACT_TANH>), dim3(grid), dim3(64 * is.waves), is.lds, stream, a);
  else
    hipLaunchKernelGGL((bf16_infer_kernel<kTiles, CCR_ACT_SILU>), dim3(grid), dim3(64 * is.waves), is.lds, stream, a);
}

hipError_t launchInfer(const MlpDevice& m, InferArgs& a, const double* mean, const double* sd, hipStream_t stream) {
  const MlpShape& sh = m.sh;
  if (m.precision != CCR_BF16 || sh.pad[0] != 32 || sh.layers < 2) return hipErrorInvalidValue;
  for (int l = 1; l < sh.layers; ++l)
    if (sh.size[l] != 128) return hipErrorInvalidValue;
  for (int l = 0; l < sh.layers; ++l) {
    a.w[l] = static_cast<const __bf16*>(m.w16[l]);
    a.b[l] = m.bpad[l];
  }
  for (int j = 0; j < sh.size[0]; ++j) {
    a.mean[j] = static_cast<float>(mean[j]);
    a.rsd[j] = static_cast<float>(1.0 / sd[j]);
  }
  a.layers = sh.layers;
  a.inputs = sh.size[0];
  a.outputs = sh.size[sh.layers];
  const InferShape is = inferShape(sh);
  if (is.tiles == 0) return hipErrorInvalidValue;
  a.w0stride = is.w0stride;
  a.wstride = is.wstride;
  const int64_t groups = (a.n + 16 * is.tiles - 1) / (16 * is.tiles);
  const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((groups + is.waves - 1) / is.waves, 4096)));
  const void* fns[4] = {reinterpret_cast<const void*>(&bf16_infer_kernel<1, CCR_ACT_TANH>),
                        reinterpret_cast<const void*>(&bf16_infer_kernel<1, CCR_ACT_SILU>),
                        reinterpret_cast<const void*>(&bf16_infer_kernel<2, CCR_ACT_TANH>),
                        reinterpret_cast<const void*>(&bf16_infer_kernel<2, CCR_ACT_SILU>)};
  for (const void* fn : fns)
    if (hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(is.lds)))
      return e;
  if (is.tiles == 2) inferLaunch<2>(a, sh.act, is, grid, stream);
  else inferLaunch<1>(a, sh.act, is, grid, stream);
  return hipGetLastError();
}

hipError_t last() { return hipGetLastError(); }

unsigned blocksFor(int64_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

hipError_t mlpInfer(const MlpDevice& m, const float* x, int64_t n, const double* mean, const double* sd, float* y,
                    hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (!y) return hipErrorInvalidValue;
  InferArgs a{};
  a.x = x;
  a.y = y;
  a.n = n;
  return launchInfer(m, a, mean, sd, stream);
}

hipError_t mlpInferRows(const MlpDevice& m, const float* x, int64_t n, const double* mean, const double* sd,
                        int64_t paths, int32_t urows, const int32_t* row_map, const InferOut* outs, double* rows,
                        hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (!rows || !row_map || !outs || paths <= 0 || urows <= 0 || n > paths * urows) return hipErrorInvalidValue;
  InferArgs a{};
  a.x = x;
  a.n = n;
  a.rows = rows;
  a.row_map = row_map;
  a.outs = outs;
  a.paths = paths;
  a.urows = urows;
  return launchInfer(m, a, mean, sd, stream);
}

hipError_t mlpLoadInputs(const MlpDevice& m, const float* xs, hipStream_t stream) {
  if (m.precision != CCR_BF16) return hipSuccess;
  hipLaunchKernelGGL(bf16_inputs_kernel, dim3(blocksFor(m.n_pad * m.sh.pad[0], 256)), dim3(256), 0, stream, xs,
                     m.sh.size[0], m.sh.pad[0], m.n, m.ld, static_cast<__bf16*>(m.a_sm[0]),
                     static_cast<__bf16*>(m.a_fm[0]));
  return last();
}

hipError_t mlpRefreshBf16(const MlpDevice& m, hipStream_t stream) {
  if (m.precision != CCR_BF16) return hipSuccess;
  for (int l = 0; l < m.sh.layers; ++l)
    hipLaunchKernelGGL(bf16_refresh_kernel, dim3(blocksFor(static_cast<int64_t>(m.sh.pad[l]) * m.sh.pad[l + 1], 256)),
                       dim3(256), 0, stream, m.params, m.sh.offW[l], m.sh.offB[l], m.sh.size[l], m.sh.size[l + 1],
                       m.sh.pad[l], m.sh.pad[l + 1], static_cast<__bf16*>(m.w16[l]), static_cast<__bf16*>(m.wt16[l]),
                       m.bpad[l]);
  return last();
}

hipError_t mlpForward(const MlpDevice& m, hipStream_t stream) {
  const MlpShape& sh = m.sh;
  const int L = sh.layers;
  if (m.precision == CCR_FP32) {
    for (int l = 0; l < L; ++l)
      hipLaunchKernelGGL(f32_fwd_kernel, dim3(blocksFor(m.n * sh.size[l + 1], 256)), dim3(256), 0, stream,
                         static_cast<const float*>(m.a_sm[l]), sh.size[l], m.params + sh.offW[l], m.params + sh.offB[l],
                         sh.size[l + 1], m.z[l + 1], static_cast<float*>(m.a_sm[l + 1]), m.n, sh.act, l + 1 == L ? 1 : 0);
    return last();
  }
  const dim3 grid(static_cast<unsigned>(m.n_pad / kMlpRows)), blk(256);
  for (int l = 0; l < L; ++l) {
    const __bf16* ain = static_cast<const __bf16*>(m.a_sm[l]);
    const __bf16* w = static_cast<const __bf16*>(m.w16[l]);
    __bf16* asm_ = static_cast<__bf16*>(m.a_sm[l + 1]);
    __bf16* afm = static_cast<__bf16*>(m.a_fm[l + 1]);
    const bool lastL = l + 1 == L;
    const int pi = sh.pad[l], po = sh.pad[l + 1];
    if (pi == 32 && po == 128 && !lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<32, 128, false>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 128 && po == 128 && !lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<128, 128, false>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 128 && po == 32 && lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<128, 32, true>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 128 && po == 64 && lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<128, 64, true>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 128 && po == 96 && lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<128, 96, true>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 128 && po == 128 && lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<128, 128, true>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else if (pi == 32 && po == 32 && lastL)
      hipLaunchKernelGGL((bf16_fwd_kernel<32, 32, true>), grid, blk, 0, stream, ain, w, m.bpad[l], m.z[l + 1], asm_, afm, m.ld, sh.act);
    else
      return hipErrorInvalidValue;
  }
  return last();
}

hipError_t mlpLossGrad(const MlpDevice& m, hipStream_t stream) {
  const MlpShape& sh = m.sh;
  const int L = sh.layers, nout = sh.size[L];
  const unsigned blocks = blocksFor(m.n * nout, 256);
  if (m.precision == CCR_FP32) {
    hipLaunchKernelGGL(f32_loss_kernel, dim3(blocks), dim3(256), 0, stream, static_cast<const float*>(m.a_sm[L]), m.ys,
                       static_cast<float*>(m.d_sm[L]), m.n, nout, m.loss_part);
    return last();
  }
  hipLaunchKernelGGL(bf16_loss_kernel, dim3(blocks), dim3(256), 0, stream, m.z[L], sh.pad[L], m.ys, nout, m.n, m.ld,
                     static_cast<__bf16*>(m.d_sm[L]), static_cast<__bf16*>(m.d_fm[L]), m.loss_part);
  return last();
}

hipError_t mlpBackward(const MlpDevice& m, hipStream_t stream) {
  const MlpShape& sh = m.sh;
  const int L = sh.layers;
  if (m.precision == CCR_FP32) {
    for (int l = L - 1; l >= 1; --l)
      hipLaunchKernelGGL(f32_bwd_kernel, dim3(blocksFor(m.n * sh.size[l], 256)), dim3(256), 0, stream,
                         static_cast<const float*>(m.d_sm[l + 1]), sh.size[l + 1], m.params + sh.offW[l], sh.size[l], m.z[l],
                         static_cast<float*>(m.d_sm[l]), m.n, sh.act);
    const float* const* acts = m.tables;
    const float* const* deltas = m.tables + (L + 1);
    hipLaunchKernelGGL(f32_wgrad_kernel, dim3(blocksFor(sh.count, 256), static_cast<unsigned>(m.nchunks)), dim3(256), 0,
                       stream, sh, acts, deltas, m.n, m.partial);
  } else {
    const dim3 grid(static_cast<unsigned>(m.n_pad / kMlpRows)), blk(256);
    for (int l = L - 1; l >= 1; --l) {
      const __bf16* dn = static_cast<const __bf16*>(m.d_sm[l + 1]);
      const __bf16* wt = static_cast<const __bf16*>(m.wt16[l]);
      __bf16* dsm = static_cast<__bf16*>(m.d_sm[l]);
      __bf16* dfm = static_cast<__bf16*>(m.d_fm[l]);
      const int pi = sh.pad[l], po = sh.pad[l + 1];
      if (pi == 128 && po == 32)
        hipLaunchKernelGGL((bf16_bwd_kernel<128, 32>), grid, blk, 0, stream, dn, wt, m.z[l], dsm, dfm, m.ld, sh.act);
      else if (pi == 128 && po == 64)
        hipLaunchKernelGGL((bf16_bwd_kernel<128, 64>), grid, blk, 0, stream, dn, wt, m.z[l], dsm, dfm, m.ld, sh.act);
      else if (pi == 128 && po == 96)
        hipLaunchKernelGGL((bf16_bwd_kernel<128, 96>), grid, blk, 0, stream, dn, wt, m.z[l], dsm, dfm, m.ld, sh.act);
      else if (pi == 128 && po == 128)
        hipLaunchKernelGGL((bf16_bwd_kernel<128, 128>), grid, blk, 0, stream, dn, wt, m.z[l], dsm, dfm, m.ld, sh.act);
      else
        return hipErrorInvalidValue;
    }
    for (int l = 0; l < L; ++l)
      hipLaunchKernelGGL(bf16_wgrad_kernel, dim3(static_cast<unsigned>((sh.pad[l + 1] / 16) * (sh.pad[l] / 16)),
                                                 static_cast<unsigned>(m.nchunks)),
                         dim3(64), 0, stream, static_cast<const __bf16*>(m.d_fm[l + 1]),
                         static_cast<const __bf16*>(m.a_fm[l]), m.n_pad, m.ld, sh.size[l], sh.size[l + 1], sh.pad[l],
                         sh.offW[l], sh.offB[l], sh.count, m.partial);
  }
  hipLaunchKernelGGL(reduce_kernel, dim3(blocksFor(sh.count, 256)), dim3(256), 0, stream, m.partial, sh.count, m.nchunks,
                     m.grad);
  return last();
}

hipError_t mlpAdam(const MlpDevice& m, float lr, float wd, float c1, float c2, hipStream_t stream) {
  hipLaunchKernelGGL(adam_kernel, dim3(blocksFor(m.sh.count, 256)), dim3(256), 0, stream, m.sh, m.params, m.grad, m.mom,
                     m.vel, lr, wd, c1, c2);
  if (hipError_t e = last()) return e;
  return mlpRefreshBf16(m, stream);
}

}  // namespace ccrhip
