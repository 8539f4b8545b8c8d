// Shared definitions for the relaxmap_amd HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace rmx {

// Teleportation (Module.h:114-116): beta is the double nearest 1.0 - 0.15.
constexpr double kAlpha = 0.15;
constexpr double kBeta = 1.0 - 0.15;

// Module statistics and per-candidate flow sums are accumulated in 64-bit
// fixed point so that every sum is exact and independent of the order in
// which wavefronts commit: results are bit-reproducible run to run.  All
// summed quantities are probabilities (<= 1); the scale leaves headroom to
// 4.0 and resolves 2^-61 ~ 4.3e-19.
constexpr double kFxScale = 2305843009213693952.0;   // 2^61
constexpr double kFxInv = 1.0 / 2305843009213693952.0;

__host__ __device__ inline long long fx_of(double x) { return (long long)__builtin_rint(x * kFxScale); }
__host__ __device__ inline double fx_to(long long v) { return (double)v * kFxInv; }

// Fixed-point statistics of a module (P flow, T teleport weight, D dangling
// flow, X flow on arcs leaving the module) or of a unit (p, tau, d, w out-flow):
// one 32-byte record, so a gather touches a single cache line.
struct alignas(32) Stat4 {
  long long p, t, d, x;
};

constexpr int kWave = 64;
constexpr uint32_t kEmptyKey = 0xffffffffu;

struct HipError : std::runtime_error {
  explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define RMX_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw ::rmx::HipError(std::string(#expr) + ": " + hipGetErrorString(_e) +    \
                            " (" __FILE__ ":" + std::to_string(__LINE__) + ")");   \
  } while (0)

#define RMX_LAUNCH_CHECK() RMX_HIP(hipGetLastError())

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// --------------------------------------------------------------- device ---
__device__ inline double plogp(double x) { return x > 0.0 ? x * log(x) : 0.0; }

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block reduction in a fixed order (deterministic for a fixed launch shape).
template <typename T, int BLOCK>
__device__ inline T block_sum(T v, T *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) r += lds[i];
  }
  __syncthreads();
  return r;   // valid in thread 0
}

__device__ inline uint32_t hash_u32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace rmx
