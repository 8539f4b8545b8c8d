// Device helpers shared by the codelength and move-sweep kernels.
#pragma once

#include "context.h"

namespace rmx {

// Arc weights in fixed point, one array per direction and level: at the node
// level the FP64 (or, in mixed precision, FP32) flow of each arc converted
// once by the steady state, at a contracted level the exact fixed-point sums.
struct ArcW {
  const long long *wx;
  __device__ long long fxw(uint32_t k) const { return wx[k]; }
};

ArcW arcw_out(const Ctx &c, const Level &L);
ArcW arcw_in(const Ctx &c, const Level &L);

// Exit probability of a module (Module.cpp:2978-2991):
//   q = (alpha P + beta D)(1 - T) + beta X
__device__ inline double module_exit(long long P, long long T, long long D, long long X) {
  return (kAlpha * fx_to(P) + kBeta * fx_to(D)) * (1.0 - fx_to(T)) + kBeta * fx_to(X);
}

}  // namespace rmx
