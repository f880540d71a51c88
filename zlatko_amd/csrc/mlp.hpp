// mlp.hpp -- the MLP surrogate on the GPU: shapes, device buffers and the launchers of mlp.hip.
//
// Network: inputs -> hidden_layers x hidden (tanh or SiLU) -> outputs (linear), the reference's
// ccr::ai::Mlp (include/ccr/ai/surrogate.hpp:44-50) generalised in depth and activation. Training
// is the reference's fitMlp (src/ai/surrogate.cpp:143-245): MSE on standardised rows, full-batch
// Adam with weight decay on weights only. Two precisions:
//   FP32: vector-unit kernels for any width <= 128 (parity with the FP64 restatement);
//   BF16: matrix-core (v_mfma_f32_16x16x32_bf16) kernels for hidden = 128, FP32 accumulation and
//         FP32 master weights / Adam moments.
// Layouts (n_pad = samples rounded up to kMlpRows; feature widths padded to kMlpPad):
//   activations and deltas, sample-major [n_pad][f] (the A operand of the forward and backward-data
//   products) and feature-major [f][n_pad] (both operands of the weight-gradient product);
//   weights W_l [f_out][f_in] and their transposes, BF16, zero-padded.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/ccr_hip.h"

namespace ccrhip {

constexpr int kMlpMaxLayers = 6;   // weight layers (hidden_layers + 1)
constexpr int kMlpMaxWidth = 128;
constexpr int kMlpRows = 64;       // sample padding (four 16-row MFMA tiles per workgroup)
constexpr int kMlpPad = 32;        // feature padding (one MFMA K step)
constexpr int kMlpChunk = 4096;    // samples per weight-gradient partial (fixed summation order)

struct MlpShape {
  int layers = 0;                  // weight layers
  int act = CCR_ACT_TANH;
  int size[kMlpMaxLayers + 1];     // feature widths
  int pad[kMlpMaxLayers + 1];      // padded widths (BF16 path)
  int64_t offW[kMlpMaxLayers], offB[kMlpMaxLayers];  // into the flattened FP32 parameters
  int64_t count = 0;
};
MlpShape mlpShape(const ccr_mlp_desc& d);

// Device state of one network and one sample set.
struct MlpDevice {
  MlpShape sh;
  int precision = CCR_FP32;
  int64_t n = 0, n_pad = 0, nchunks = 0;
  int64_t ld = 0;            // row length of the feature-major arrays (the full sample count, padded)
  float* params = nullptr;   // FP32 master parameters [count]
  float* grad = nullptr;     // [count]
  float* mom = nullptr;
  float* vel = nullptr;
  float* partial = nullptr;  // [nchunks][count] weight-gradient partials
  float* xs = nullptr;       // FP32 path: standardised inputs [n][in]; BF16: unused
  float* ys = nullptr;       // standardised targets [n][out]
  float* loss_part = nullptr;  // per-block squared-error partials
  int loss_blocks = 0;
  // Per layer boundary l = 0..layers: activations (FP32 path: a[l] sample-major FP32, z[l]);
  // BF16 path: a_sm / a_fm BF16, z FP32 sample-major (padded widths).
  void* a_sm[kMlpMaxLayers + 1] = {};
  void* a_fm[kMlpMaxLayers + 1] = {};
  float* z[kMlpMaxLayers + 1] = {};
  void* d_sm[kMlpMaxLayers + 1] = {};
  void* d_fm[kMlpMaxLayers + 1] = {};
  void* w16[kMlpMaxLayers] = {};   // BF16 W_l [pad_out][pad_in]
  void* wt16[kMlpMaxLayers] = {};  // BF16 W_l^T [pad_in][pad_out]
  float* bpad[kMlpMaxLayers] = {}; // FP32 biases, padded
  float** tables = nullptr;        // FP32 path: device copy of a_sm[0..layers] then d_sm[0..layers]
};

// Launchers (mlp.hip). All work on `stream`; none allocates.
hipError_t mlpForward(const MlpDevice& m, hipStream_t stream);          // fills a/z/out
hipError_t mlpLossGrad(const MlpDevice& m, hipStream_t stream);         // output deltas + loss partials
hipError_t mlpBackward(const MlpDevice& m, hipStream_t stream);         // deltas, gradient partials, gradient
hipError_t mlpAdam(const MlpDevice& m, float lr, float wd, float c1, float c2, hipStream_t stream);
hipError_t mlpRefreshBf16(const MlpDevice& m, hipStream_t stream);      // BF16 weight copies from params
// BF16 inputs from FP32 standardised rows (the BF16 path's layer-0 activations).
hipError_t mlpLoadInputs(const MlpDevice& m, const float* xs, hipStream_t stream);
// Fused BF16 inference with m's weights (hidden = 128, inputs <= 32, outputs <= 128): raw rows
// x [n][inputs], standardised on load with (mean, sd); standardised outputs y [n][outputs]. Needs no
// per-sample buffers.
hipError_t mlpInfer(const MlpDevice& m, const float* x, int64_t n, const double* mean, const double* sd, float* y,
                    hipStream_t stream);
// Destination of one network output in mlpInferRows: rows base + r of a [row][path] FP64 array,
// de-standardised with (mean, sd).
struct InferOut {
  int64_t base;
  double mean, sd;
};
// mlpInfer with the outputs written straight into value rows: sample s = u * paths + p (u < urows);
// output o goes to rows[(outs[o].base + row_map[o * urows + u]) * paths + p] where row_map >= 0.
hipError_t mlpInferRows(const MlpDevice& m, const float* x, int64_t n, const double* mean, const double* sd,
                        int64_t paths, int32_t urows, const int32_t* row_map, const InferOut* outs, double* rows,
                        hipStream_t stream);

}  // namespace ccrhip
