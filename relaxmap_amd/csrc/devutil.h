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

// Rows of a CSR with more than kBigRow arcs, for kernels that give every row
// a small group of lanes: such a kernel skips the big rows, and a companion
// kernel takes their arcs in chunks of kBigRow, one workgroup per chunk, so
// that a row of 10^5 arcs does not hold one group for the whole launch.
// ctr[0] counts the chunks; chunk i covers arcs [cbeg[i], min(cbeg[i] +
// kBigRow, row end)) of row crow[i].  At most 2 m / kBigRow + 1 chunks;
// *arcs counts the arcs of the big rows (for the profiling mode's bytes).
constexpr uint32_t kBigRow = 1024;

struct BigRows {
  uint32_t *ctr, *crow, *cbeg;
  unsigned long long *arcs;
};

inline int64_t big_rows_cap(int64_t m) { return 2 * (m / kBigRow) + 1; }

// fills B.ctr and plans the chunks of the rows of (ptr, n); the arrays come
// from the caller (big_rows_cap(m) chunks, one counter)
void plan_big_rows(const uint32_t *ptr, int32_t n, const BigRows &B, hipStream_t s);

// grid of a chunk kernel: it strides over the chunks (their count is read on
// the device)
inline unsigned big_rows_grid(int64_t m) { return (unsigned)std::min<int64_t>(big_rows_cap(m), 2048); }

}  // namespace rmx
