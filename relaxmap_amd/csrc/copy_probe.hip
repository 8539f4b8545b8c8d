// Calibration probe for HBM traffic counters: a plain device copy of a known
// size (1 GiB read, 1 GiB written per launch).  rocprofv3's FETCH_SIZE and
// WRITE_SIZE for this kernel give the factors that turn those counters into
// bytes for the library's kernels.
//   rocprofv3 --pmc FETCH_SIZE -- relaxmap_amd/copy_probe
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy_probe(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long s = (long)gridDim.x * blockDim.x;
  for (; i < n; i += s) b[i] = a[i];
}

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
      return 1;                                                               \
    }                                                                         \
  } while (0)

int main() {
  const long bytes = 1L << 30;
  const long n = bytes / (long)sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 1, bytes));
  CHECK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  for (int r = 0; r < 5; ++r) {
    CHECK(hipEventRecord(e0));
    k_copy_probe<<<4096, 256>>>(a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy of %ld bytes: %.3f ms, %.1f GB/s read + write\n", bytes, ms, 2.0 * bytes / ms / 1e6);
  }
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return 0;
}
