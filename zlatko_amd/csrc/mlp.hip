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

// Output deltas 2 (y - t) / (n out) in BF16 (sample- and feature-major) and squared-error sums.
__global__ void __launch_bounds__(256) bf16_loss_kernel(const float* out, int po, const float* ys, int nout, int64_t n,
                                                        int64_t n_pad, __bf16* d_sm, __bf16* d_fm, float* loss_part) {
  __shared__ double lds[4];
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (idx < n * nout) {
    const int64_t s = idx / nout;
    const int k = static_cast<int>(idx % nout);
    const float e = out[s * po + k] - ys[idx];
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

// Fused inference (hidden = 128): every layer of a 16-sample tile per wavefront, weights staged
// in LDS, each layer's output turned into the next layer's A operand through a per-wavefront LDS
// tile. Inputs x [n][inputs] (raw) are standardised on load with (mean, sd); writes the
// standardised outputs y [n][outputs]. Grid-stride over tiles. The wavefronts per workgroup
// (blockDim.x / 64, at most kInferWaves) are chosen at launch so that the weights and one
// activation tile per wavefront fit the workgroup's LDS.
constexpr int kInferWaves = 16;
constexpr int kActStride = 128 + 8;  // BF16 elements per activation row in LDS (padding vs bank conflicts)
constexpr size_t kInferLdsBytes = 160 * 1024;  // LDS of one gfx950 workgroup
struct InferArgs {
  const __bf16* w[kMlpMaxLayers];
  const float* b[kMlpMaxLayers];
  const float* x;
  float* y;
  float mean[32], sd[32];
  int64_t n;
  int layers, inputs, outputs, act;
};

__device__ inline void stageWeights(__bf16* dst, const __bf16* src, int elems) {
  for (int i = static_cast<int>(threadIdx.x) * 8; i < elems; i += static_cast<int>(blockDim.x) * 8)
    *reinterpret_cast<bf16x8*>(dst + i) = load8(src + i);
}

__global__ void __launch_bounds__(64 * kInferWaves) bf16_infer_kernel(InferArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __bf16* wl = reinterpret_cast<__bf16*>(lds_raw);
  // Layers in order: first 128 x 32, middle 128 x 128, last 32 x 128 (32 x 32 alone).
  int wtotal = 0;
  for (int l = 0; l < a.layers; ++l) {
    const int fi = l == 0 ? 32 : 128, fo = l + 1 == a.layers ? 32 : 128;
    stageWeights(wl + wtotal, a.w[l], fi * fo);
    wtotal += fi * fo;
  }
  __bf16* act = wl + wtotal + (static_cast<int>(threadIdx.x) >> 6) * 16 * kActStride;
  __syncthreads();
  const int lane = static_cast<int>(threadIdx.x) & 63, r = lane & 15, g = lane >> 4;
  const int waves = static_cast<int>(blockDim.x) >> 6;
  const int64_t tiles = (a.n + 15) / 16;
  for (int64_t tile = static_cast<int64_t>(blockIdx.x) * waves + (threadIdx.x >> 6); tile < tiles;
       tile += static_cast<int64_t>(gridDim.x) * waves) {
    const int64_t s0 = tile * 16;
    // Layer-0 A operand: sample s0 + r, features 8 g .. 8 g + 7.
    bf16x8 a0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int j = 8 * g + i;
      const int64_t s = s0 + r;
      float v = 0.0f;
      if (j < a.inputs && s < a.n) v = (a.x[s * a.inputs + j] - a.mean[j]) / a.sd[j];
      a0[i] = static_cast<__bf16>(v);
    }
    int wo = 0;
    for (int l = 0; l < a.layers; ++l) {
      const bool lastL = l + 1 == a.layers;
      const int fi = l == 0 ? 32 : 128, nt = lastL ? 2 : 8;
      const __bf16* w = wl + wo;
      wo += fi * nt * 16;
      bf16x8 afr[4];
      if (l == 0) {
        afr[0] = a0;
      } else {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) afr[ks] = *reinterpret_cast<const bf16x8*>(act + r * kActStride + ks * 32 + 8 * g);
      }
      for (int t = 0; t < nt; ++t) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (l == 0) {
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[0], *reinterpret_cast<const bf16x8*>(w + (t * 16 + r) * fi + 8 * g),
                                                        acc, 0, 0, 0);
        } else {
#pragma unroll
          for (int ks = 0; ks < 4; ++ks)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                afr[ks], *reinterpret_cast<const bf16x8*>(w + (t * 16 + r) * fi + ks * 32 + 8 * g), acc, 0, 0, 0);
        }
        const int o = t * 16 + r;
        const float bo = a.b[l][o];
        if (lastL) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int64_t s = s0 + 4 * g + i;
            if (o < a.outputs && s < a.n) a.y[s * a.outputs + o] = acc[i] + bo;
          }
        } else {
          // This wavefront's tile rows 4 g + i, column o (all lanes have read the previous
          // layer's rows into afr before any lane writes: wavefront-synchronous).
#pragma unroll
          for (int i = 0; i < 4; ++i) act[(4 * g + i) * kActStride + o] = static_cast<__bf16>(actInfer(a.act, acc[i] + bo));
        }
      }
    }
  }
}

hipError_t last() { return hipGetLastError(); }

unsigned blocksFor(int64_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

hipError_t mlpInfer(const MlpDevice& m, const float* x, int64_t n, const double* mean, const double* sd, float* y,
                    hipStream_t stream) {
  const MlpShape& sh = m.sh;
  if (m.precision != CCR_BF16 || sh.pad[0] != 32 || sh.pad[sh.layers] != 32) return hipErrorInvalidValue;
  for (int l = 1; l < sh.layers; ++l)
    if (sh.size[l] != 128) return hipErrorInvalidValue;
  InferArgs a{};
  for (int l = 0; l < sh.layers; ++l) {
    a.w[l] = static_cast<const __bf16*>(m.w16[l]);
    a.b[l] = m.bpad[l];
  }
  for (int j = 0; j < sh.size[0]; ++j) {
    a.mean[j] = static_cast<float>(mean[j]);
    a.sd[j] = static_cast<float>(sd[j]);
  }
  a.x = x;
  a.y = y;
  a.n = n;
  a.layers = sh.layers;
  a.inputs = sh.size[0];
  a.outputs = sh.size[sh.layers];
  a.act = sh.act;
  int welems = 0;
  for (int l = 0; l < sh.layers; ++l) welems += sh.pad[l] * sh.pad[l + 1];
  // As many wavefronts (16 at most) as leave room for one activation tile each beside the weights:
  // 16 up to 3 hidden layers, 3 at 5 hidden layers.
  const size_t wbytes = static_cast<size_t>(welems) * 2, tile = static_cast<size_t>(16) * kActStride * 2;
  if (wbytes + tile > kInferLdsBytes) return hipErrorInvalidValue;
  const int waves = static_cast<int>(std::min<size_t>(kInferWaves, (kInferLdsBytes - wbytes) / tile));
  const size_t lds = wbytes + static_cast<size_t>(waves) * tile;
  const int64_t tiles = (n + 15) / 16;
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>((tiles + waves - 1) / waves, 2048));
  if (hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&bf16_infer_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)))
    return e;
  hipLaunchKernelGGL(bf16_infer_kernel, dim3(grid), dim3(64 * waves), lds, stream, a);
  return last();
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
