// Network::calculateCodeLength (Module.cpp:2966-3002) on caller-owned device
// arrays, with no device allocation: every temporary of the computation (the
// module sort, segment starts, per-node exit flows, per-module exits,
// reduction partials, the sort's look-back words) lives in one file-scope
// device array, `scratch`, sized for at most kMaxNodes nodes.
//
// Each GPU holds its own copy of `scratch` for the life of the process.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in for the device symbol): the first call on a device, with that
// device current, looks up the device address with hipGetSymbolAddress and
// keeps it per device; every kernel argument and copy uses that address.
// Calls that overlap on one GPU share that GPU's array and race; a call
// leaves its contents behind, and everything a call needs zeroed is cleared
// by a fill kernel at its start.
#include <cstring>
#include <mutex>
#include <string>

#include "devutil.h"
#include "../../include/relaxmap.h"

namespace rmx {
namespace {

constexpr int32_t kMaxNodes = RMX_NOALLOC_MAX_NODES;
constexpr size_t kScratchBytes = (size_t)56 << 20;   // 52.4 MiB at kMaxNodes
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Layout of the array for n nodes.
struct Carve {
  size_t key, node, flag, pos, start, total, ext, segq, segqp, partial, scal, check, arena, arena_bytes, end;
};

Carve carve(int32_t n) {
  Carve c;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o += align256(b);
    return at;
  };
  c.key = take(sizeof(uint32_t) * n);
  c.node = take(sizeof(uint32_t) * n);
  c.flag = take(sizeof(uint32_t) * n);
  c.pos = take(sizeof(uint32_t) * n);
  c.start = take(sizeof(uint32_t) * n);
  c.total = take(sizeof(uint32_t));
  c.ext = take(sizeof(double) * n);
  c.segq = take(sizeof(double) * n);
  c.segqp = take(sizeof(double) * n);
  c.partial = take(sizeof(double) * p5_partial_doubles());
  c.scal = take(sizeof(double) * 8);
  c.check = take(sizeof(uint32_t));
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
  c.arena_bytes = std::max(radix_sort_scratch_bytes(n, 4, 4, bits), exclusive_scan_scratch_bytes(n));
  c.arena = take(c.arena_bytes);
  c.end = o;
  return c;
}

// Input checks that keep every later access in bounds: out_ptr starts at 0,
// ends at m and never decreases; every arc head and every module index lies
// in [0, n).  Bit 1, 2, 4 of *bad respectively.
__global__ void k_check_inputs(int32_t n, int64_t m, const uint32_t *__restrict__ optr,
                               const uint32_t *__restrict__ oidx, const int32_t *__restrict__ mod,
                               uint32_t *__restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t flags = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
    if (i == 0 && optr[0] != 0u) flags |= 1u;
    if (i == n && (int64_t)optr[n] != m) flags |= 1u;
    if (i < n) {
      if (optr[i + 1] < optr[i]) flags |= 1u;
      if ((uint32_t)mod[i] >= (uint32_t)n) flags |= 4u;
    }
  }
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride)
    if (oidx[k] >= (uint32_t)n) flags |= 2u;
  if (flags) atomicOr(bad, flags);
}

void *device_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    RMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

void put_error(char *err, int err_len, const std::string &msg) {
  if (!err || err_len <= 0) return;
  const size_t k = std::min(msg.size(), (size_t)err_len - 1);
  std::memcpy(err, msg.data(), k);
  err[k] = '\0';
}

}  // namespace
}  // namespace rmx

extern "C" int rmx_calculate_codelength_noalloc(int device, void *stream, int32_t n, int64_t m,
                                                const uint32_t *d_out_ptr, const uint32_t *d_out_idx,
                                                const double *d_flow, const double *d_p, const double *d_tau,
                                                const uint8_t *d_dangling, const int32_t *d_module,
                                                double *codelength, char *err, int err_len) {
  using namespace rmx;
  put_error(err, err_len, "");
  if (n < 1 || n > kMaxNodes) {
    put_error(err, err_len,
              "n = " + std::to_string(n) + " nodes is outside [1, " + std::to_string(kMaxNodes) +
                  "], the size of the static scratch array");
    return RMX_EINVAL;
  }
  if (m < 0 || m > (int64_t)INT32_MAX) {
    put_error(err, err_len, "m = " + std::to_string(m) + " arcs is outside [0, 2^31)");
    return RMX_EINVAL;
  }
  if (!d_out_ptr || !d_p || !d_tau || !d_dangling || !d_module || !codelength || (m > 0 && (!d_out_idx || !d_flow))) {
    put_error(err, err_len, "a required pointer is NULL");
    return RMX_EINVAL;
  }
  if (device < 0 || device >= kMaxDevices) {
    put_error(err, err_len, "device index out of range");
    return RMX_EINVAL;
  }
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  int rc = RMX_OK;
  try {
    int count = 0;
    RMX_HIP(hipGetDeviceCount(&count));
    if (device >= count) throw HipError("no such device");
    RMX_HIP(hipSetDevice(device));
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)device_scratch(device);
    const Carve cv = carve(n);
    if (cv.end > kScratchBytes) throw HipError("scratch layout exceeds the static array");
    uint32_t *check = (uint32_t *)(base + cv.check);
    fill_u32(check, 0u, 1, s);
    const unsigned check_grid = std::min<unsigned>(grid_for(std::max<int64_t>((int64_t)n + 1, m), 256), 4096u);
    k_check_inputs<<<check_grid, 256, 0, s>>>(n, m, d_out_ptr, d_out_idx, d_module, check);
    RMX_LAUNCH_CHECK();
    uint32_t h_check = 0;
    RMX_HIP(hipMemcpyAsync(&h_check, check, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RMX_HIP(hipStreamSynchronize(s));
    if (h_check) {
      std::string why;
      if (h_check & 1u) why += " out_ptr is not a CSR row index of m arcs;";
      if (h_check & 2u) why += " an arc head is outside [0, n);";
      if (h_check & 4u) why += " a module index is outside [0, n);";
      put_error(err, err_len, "rejected input:" + why);
      rc = RMX_EINVAL;
    } else {
      P5Work w;
      w.key = (uint32_t *)(base + cv.key);
      w.node = (uint32_t *)(base + cv.node);
      w.flag = (uint32_t *)(base + cv.flag);
      w.pos = (uint32_t *)(base + cv.pos);
      w.start = (uint32_t *)(base + cv.start);
      w.total = (uint32_t *)(base + cv.total);
      w.ext = (double *)(base + cv.ext);
      w.segq = (double *)(base + cv.segq);
      w.segqp = (double *)(base + cv.segqp);
      w.partial = (double *)(base + cv.partial);
      w.scal = (double *)(base + cv.scal);
      Scratch arena;
      arena.ptr = base + cv.arena;
      arena.bytes = cv.arena_bytes;
      arena.fixed = true;
      p5_launch(n, d_out_ptr, d_out_idx, d_flow, (const uint32_t *)d_module, d_p, d_tau,
                d_dangling, 0.0, true, w, arena, s);
      double h[8];
      RMX_HIP(hipMemcpyAsync(h, w.scal, sizeof(h), hipMemcpyDeviceToHost, s));
      RMX_HIP(hipStreamSynchronize(s));
      *codelength = h[3];
    }
  } catch (const std::exception &e) {
    put_error(err, err_len, e.what());
    rc = RMX_EDEVICE;
  }
  if (prev >= 0) (void)hipSetDevice(prev);
  return rc;
}
