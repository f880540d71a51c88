// launch.hpp -- kernel arguments and host-side launchers of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace ccrhip {

template <typename Real>
struct PathArgs {
  const GlobalDev* g;
  const SetDev* sets;
  const int32_t* block_set;
  const Real* stream;
  const Real* steps;
  const DateRec* dates;
  const Real* date_reals;
  const Real* bonds;
  const Real* fixed;
  const int32_t* fixstore;
  Real* rows;
  double* partials;
  int32_t max_rep;
  int32_t ring;    // LDS rows per look-back ring
  int32_t slots;   // LDS rows of fixing states
};

template <typename Real>
hipError_t launchPath(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream);
hipError_t launchReduce(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                        double inv_paths, double* out, hipStream_t stream);
template <typename Real>
hipError_t launchPfe(const SetDev* sets, const Real* rows, int nsets, int max_rep, int num_paths, double level,
                     double* out, hipStream_t stream);

}  // namespace ccrhip
