// launch.hpp -- kernel arguments and host-side launchers of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

// Plan tables are read through the constant address space on the device: AMDGPU treats such
// loads as invariant and issues them as scalar (s_load) loads when the address is uniform, so
// one load feeds all 64 lanes. Host code (the CPU emulator) sees plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define CCR_CONST __attribute__((address_space(4)))
#else
#define CCR_CONST
#endif

namespace ccrhip {

template <class T>
using cptr = const CCR_CONST T*;
template <class T>
__host__ __device__ inline cptr<T> toConst(const T* p) {
  return (cptr<T>)p;  // address-space cast (a C-style cast is the form clang accepts)
}

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
  const double* poly;  // bond-polynomial coefficients (FP64 in both precisions)
  Real* rows;
  double* partials;
  int32_t max_rep;
  int32_t ring;        // LDS rows of the margin-call value ring
  int32_t price_rows;  // LDS rows of the collateral price ring (0 for cash collateral)
  int32_t slots;       // LDS rows of fixing states
  int32_t nrates, nassets, ncredits;  // model shape: selects the kernel specialisation
};

template <typename Real>
hipError_t launchPath(const PathArgs<Real>& args, int blocks, int block, hipStream_t stream);
hipError_t launchReduce(const SetDev* sets, const double* partials, int nsets, int max_rep, int waves_per_set,
                        double inv_paths, double* out, hipStream_t stream);
template <typename Real>
hipError_t launchPfe(const SetDev* sets, const Real* rows, int nsets, int max_rep, int num_paths, double level,
                     double* out, hipStream_t stream);

}  // namespace ccrhip
