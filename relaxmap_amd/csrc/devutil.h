// Device helpers shared by the codelength and move-sweep kernels.
#pragma once

#include "context.h"

namespace rmx {

// Arc weight in fixed point: level 0 converts the FP64 (or, in mixed
// precision, FP32) flow of the arc; level 1 stores exact fixed-point sums.
struct ArcW {
  const double *w;
  const float *wf;
  const long long *wx;
  int mode;   // 0: FP64 flows, 1: FP32 flows, 2: fixed-point sums
  __device__ long long fxw(uint32_t k) const {
    return mode == 2 ? wx[k] : (mode == 1 ? fx_of((double)wf[k]) : fx_of(w[k]));
  }
};

ArcW arcw_out(const Ctx &c, const Level &L);
ArcW arcw_in(const Ctx &c, const Level &L);

// Exit probability of a module (Module.cpp:2978-2991):
//   q = (alpha P + beta D)(1 - T) + beta X
__device__ inline double module_exit(long long P, long long T, long long D, long long X) {
  return (kAlpha * fx_to(P) + kBeta * fx_to(D)) * (1.0 - fx_to(T)) + kBeta * fx_to(X);
}

}  // namespace rmx
