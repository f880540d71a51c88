// mlp_host.hpp -- host side of the MLP surrogate: shape checks, the reference's initialisation and
// standardisers, and MlpRun, which owns the device buffers of one network and one sample set.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mlp.hpp"

namespace ccrhip {

// Standardiser::fit (src/ai/surrogate.cpp:64-78): mean and population sd per column, 0 -> 1.
void standardiserFit(const double* rows, int64_t n, int dim, double* mean, double* sd);
// fitMlp's initialisation (surrogate.cpp:153-162): Mulberry32 + Box-Muller N(0, 1 / fan_in).
void mlpInit(const MlpShape& sh, uint32_t seed, double* params);

class MlpRun {
 public:
  // Validates the shape for the precision and allocates every buffer for n samples.
  MlpRun(const ccr_mlp_desc& d, int precision, int64_t n, hipStream_t stream);
  ~MlpRun();
  MlpRun(const MlpRun&) = delete;
  MlpRun& operator=(const MlpRun&) = delete;

  void setParams(const double* params);        // FP64 host -> FP32 master (+ BF16 copies)
  void getParams(double* params) const;
  void setInputs(const double* xs);             // standardised rows [n][inputs]
  void setTargets(const double* ys);            // standardised rows [n][outputs]
  double lossAndGradient(double* grad);         // forward + backward at the current parameters
  // Adam: epochs passes over the samples, in steps of `batch` consecutive samples (0: full
  // batch, the reference's scheme); the step counter continues across calls.
  void train(int epochs, double lr, double wd, int64_t batch = 0);
  double loss();                                // forward + loss at the current parameters
  void outputs(double* ys) const;               // standardised network outputs [n][outputs] (after a forward)
  void forward();
  const MlpDevice& device() const { return m_; }

 private:
  void* alloc(size_t bytes);
  MlpDevice m_;
  hipStream_t stream_;
  std::vector<void*> bufs_;
  int epoch_ = 0;
};

}  // namespace ccrhip
