// Launch interface of the local gauge alignment kernels (gauge_local.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mepgl {

// One batched product of an image's square matrices (see k_dct_gemm).
struct GemmArgs {
  const double* A;
  long long ar, ac, abatch;
  const double* B;
  long long br, bc, bbatch;
  double* O;
  long long orow, obatch;
  int n;
  const double* lam;  // non-null: divide by lam[m] + lam[n], zero the (0, 0) mode
};

// which = 0: right-hand sides of every image; 1: apply chi to the fields.
hipError_t launch_glocal(const AuxArgs& a, int which, bool fp32, hipStream_t st);
hipError_t launch_dct_gemm(const GemmArgs& g, int F, hipStream_t st);
hipError_t launch_glocal_prefix(double* W, int N, int F, long long pitch, long long IS,
                                hipStream_t st);

}  // namespace mepgl
