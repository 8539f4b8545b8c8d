// Scans and a stable LSD radix sort written for CDNA4 wave64.
//
// Scan: reduce-then-scan over tiles of 4096 elements (256 threads x 16),
// recursing on the tile sums; no inter-block spinning.
// Radix sort: 8-bit digits; per pass an upsweep (per-tile digit histogram,
// digit-major so that one exclusive scan yields every tile's scatter base)
// and a downsweep that ranks each tile's elements stably with wave ballots
// (equal-digit peers within a wave) and wave-private digit counts.
#include "prims.h"
#include "context.h"

namespace rmx {

void *Scratch::get(size_t need) {
  if (need > bytes) {
    if (ptr) RMX_HIP(hipFree(ptr));
    size_t b = need + need / 4 + 4096;
    RMX_HIP(hipMalloc(&ptr, b));
    bytes = b;
  }
  return ptr;
}

void Profiler::begin(hipStream_t s) {
  if (!e0) {
    RMX_HIP(hipEventCreate(&e0));
    RMX_HIP(hipEventCreate(&e1));
  }
  RMX_HIP(hipEventRecord(e0, s));
}

float Profiler::stop(hipStream_t s) {
  RMX_HIP(hipEventRecord(e1, s));
  RMX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  RMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  last_ms = ms;
  return ms;
}

void Profiler::end(hipStream_t s, const std::string &name, double bytes) {
  float ms = stop(s);
  KernelStat &k = get(name);
  k.launches += 1;
  k.ms += ms;
  k.bytes += bytes;
}

void Scratch::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kScanBlock) void k_tile_sums(const uint32_t *__restrict__ in, int64_t n,
                                                          uint32_t *__restrict__ sums) {
  __shared__ uint32_t lds[kScanBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + (int64_t)i * kScanBlock + threadIdx.x;
    if (j < n) acc += in[j];
  }
  uint32_t t = block_sum<uint32_t, kScanBlock>(acc, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// Exclusive scan of one tile (blocked arrangement: thread t owns items
// [t*16, t*16+16)), offset by tile_base[blockIdx] (or 0 if null).
__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                          int64_t n, const uint32_t *__restrict__ tile_base,
                                                          uint32_t *__restrict__ total) {
  __shared__ uint32_t wsum[kScanBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    v[i] = j < n ? in[j] : 0u;
    acc += v[i];
  }
  // inclusive scan of acc across the block
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t x = acc;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  uint32_t wpre = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / kWave; ++i) {
    if (i < w) wpre += wsum[i];
    all += wsum[i];
  }
  uint32_t run = (tile_base ? tile_base[blockIdx.x] : 0u) + wpre + x - acc;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t j = base + i;
    if (j < n) out[j] = run;
    run += v[i];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *total = (tile_base ? tile_base[blockIdx.x] : 0u) + all;
}

size_t scan_tmp_elems(int64_t n) {
  size_t t = 0;
  while (n > kScanTile) {
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    t += (size_t)tiles;
    n = tiles;
  }
  return t + 1;
}

void scan_impl(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total, uint32_t *tmp, hipStream_t s) {
  if (n <= 0) {
    if (total) RMX_HIP(hipMemsetAsync(total, 0, sizeof(uint32_t), s));
    return;
  }
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    k_tile_scan<<<1, kScanBlock, 0, s>>>(in, out, n, nullptr, total);
    RMX_LAUNCH_CHECK();
    return;
  }
  uint32_t *sums = tmp;
  k_tile_sums<<<(unsigned)tiles, kScanBlock, 0, s>>>(in, n, sums);
  RMX_LAUNCH_CHECK();
  scan_impl(sums, sums, tiles, nullptr, tmp + tiles, s);
  k_tile_scan<<<(unsigned)tiles, kScanBlock, 0, s>>>(in, out, n, sums, total);
  RMX_LAUNCH_CHECK();
}

// ------------------------------------------------------------ radix sort
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kRadix = 256;

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_sort_upsweep(const K *__restrict__ keys, int64_t n, int shift,
                                                             int64_t n_tiles, uint32_t *__restrict__ counts) {
  __shared__ uint32_t hist[kRadix];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    int64_t j = base + (int64_t)i * kSortBlock + threadIdx.x;
    if (j < n) atomicAdd(&hist[(uint32_t)(keys[j] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = hist[threadIdx.x];
}

// Stable scatter of one tile.  Wave w owns the tile's elements
// [w * 1024, (w + 1) * 1024), taken 64 at a time in index order; it ranks
// each element among the wave's earlier elements of the same digit with
// ballots (equal-digit peers of a step) and a wave-private digit count in LDS,
// so no workgroup barrier is needed while ranking.  One digit-major scan
// over the per-wave counts then gives every (wave, digit) its first position,
// the tile is reordered by digit in LDS, and each digit's run is written to
// its global range with consecutive lanes on consecutive addresses.
template <typename K, typename V>
__global__ __launch_bounds__(kSortBlock) void k_sort_downsweep(const K *__restrict__ keys_in,
                                                               const V *__restrict__ vals_in,
                                                               K *__restrict__ keys_out,
                                                               V *__restrict__ vals_out, int64_t n,
                                                               int shift, int64_t n_tiles,
                                                               const uint32_t *__restrict__ offsets) {
  constexpr int W = kSortBlock / kWave;
  constexpr int PW = kSortTile / W;   // elements per wave
  __shared__ K s_key[kSortTile];
  __shared__ V s_val[kSortTile];
  __shared__ uint32_t hist[W][kRadix];   // per-wave digit counts, then per-wave first positions
  __shared__ uint32_t tile_off[kRadix];
  __shared__ uint32_t digit_start[kRadix];
  __shared__ uint32_t wsum[W];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const int64_t wbase = base + (int64_t)w * PW;
  tile_off[t] = offsets[(int64_t)t * n_tiles + blockIdx.x];
#pragma unroll
  for (int q = 0; q < W; ++q) hist[q][t] = 0;
  K key[kSortItems];
  V val[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int64_t j = wbase + (int64_t)i * kWave + lane;
    const bool valid = j < n;
    key[i] = valid ? keys_in[j] : K(0);
    val[i] = valid ? vals_in[j] : V(0);
  }
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t rank[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int64_t j = wbase + (int64_t)i * kWave + lane;
    const bool valid = j < n;
    const uint32_t d = (uint32_t)(key[i] >> shift) & 0xffu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t r = (uint32_t)__popcll(peers & lt_mask);
    const uint32_t before = hist[w][d];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();   // every lane has read its count before the leaders update it
    if (valid && r == 0) hist[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    rank[i] = before + r;
  }
  __syncthreads();
  {   // thread t holds digit t: its first position in the tile, then per wave
    uint32_t c[W];
    uint32_t tot = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      c[q] = hist[q][t];
      tot += c[q];
    }
    uint32_t x = tot;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, kWave);
      if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int q = 0; q < w; ++q) pre += wsum[q];
    uint32_t run = pre + x - tot;
    digit_start[t] = run;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      hist[q][t] = run;
      run += c[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int64_t j = wbase + (int64_t)i * kWave + lane;
    if (j < n) {
      const uint32_t d = (uint32_t)(key[i] >> shift) & 0xffu;
      const uint32_t p = hist[w][d] + rank[i];
      s_key[p] = key[i];
      s_val[p] = val[i];
    }
  }
  __syncthreads();
  const int64_t tile_n = (n - base) < kSortTile ? (n - base) : kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int idx = i * kSortBlock + t;
    if (idx < tile_n) {
      const K k = s_key[idx];
      const uint32_t d = (uint32_t)(k >> shift) & 0xffu;
      const uint32_t pos = tile_off[d] + ((uint32_t)idx - digit_start[d]);
      keys_out[pos] = k;
      vals_out[pos] = s_val[idx];
    }
  }
}

// With out_k/out_v the result is left where the last pass wrote it (the
// inputs or the scratch) and the pointers say where; otherwise it is copied
// back into keys/vals.
template <typename K, typename V>
void radix_sort_impl(K *keys, V *vals, int64_t n, int begin_bit, int end_bit, Scratch &scratch, hipStream_t s,
                     K **out_k = nullptr, V **out_v = nullptr) {
  if (out_k) *out_k = keys;
  if (out_v) *out_v = vals;
  if (n <= 1 || begin_bit >= end_bit) return;
  if (n > (int64_t)0xffffffffLL) throw HipError("radix sort: n >= 2^32");
  const int64_t n_tiles = (n + kSortTile - 1) / kSortTile;
  const int64_t n_counts = n_tiles * kRadix;
  const size_t scan_tmp = scan_tmp_elems(n_counts);
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off_k = 0;
  size_t off_v = off_k + align(sizeof(K) * n);
  size_t off_c = off_v + align(sizeof(V) * n);
  size_t off_t = off_c + align(sizeof(uint32_t) * n_counts);
  size_t total = off_t + align(sizeof(uint32_t) * scan_tmp);
  char *base = (char *)scratch.get(total);
  K *tk = (K *)(base + off_k);
  V *tv = (V *)(base + off_v);
  uint32_t *counts = (uint32_t *)(base + off_c);
  uint32_t *stmp = (uint32_t *)(base + off_t);
  K *src_k = keys, *dst_k = tk;
  V *src_v = vals, *dst_v = tv;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
    k_sort_upsweep<K><<<(unsigned)n_tiles, kSortBlock, 0, s>>>(src_k, n, shift, n_tiles, counts);
    RMX_LAUNCH_CHECK();
    scan_impl(counts, counts, n_counts, nullptr, stmp, s);
    k_sort_downsweep<K, V><<<(unsigned)n_tiles, kSortBlock, 0, s>>>(src_k, src_v, dst_k, dst_v, n, shift, n_tiles,
                                                                    counts);
    RMX_LAUNCH_CHECK();
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (out_k) {
    *out_k = src_k;
    *out_v = src_v;
  } else if (passes & 1) {
    RMX_HIP(hipMemcpyAsync(keys, src_k, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
    RMX_HIP(hipMemcpyAsync(vals, src_v, sizeof(V) * n, hipMemcpyDeviceToDevice, s));
  }
}

template <typename T>
__global__ void k_fill(T *p, T v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_iota(uint32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (uint32_t)i;
}

unsigned fill_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total, Scratch &scratch,
                        hipStream_t s) {
  uint32_t *tmp = (uint32_t *)scratch.get(sizeof(uint32_t) * scan_tmp_elems(n));
  scan_impl(in, out, n, d_total, tmp, s);
}

void radix_sort_pairs_u32(uint32_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s) {
  radix_sort_impl<uint32_t, uint32_t>(keys, vals, n, begin_bit, end_bit, scratch, s);
}

void radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s) {
  radix_sort_impl<uint64_t, uint32_t>(keys, vals, n, begin_bit, end_bit, scratch, s);
}

void radix_sort_pairs_u64_u64(uint64_t *keys, uint64_t *vals, int64_t n, int begin_bit, int end_bit,
                              Scratch &scratch, hipStream_t s, uint64_t **out_keys, uint64_t **out_vals) {
  radix_sort_impl<uint64_t, uint64_t>(keys, vals, n, begin_bit, end_bit, scratch, s, out_keys, out_vals);
}

void fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<uint32_t><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  RMX_LAUNCH_CHECK();
}
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<uint8_t><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  RMX_LAUNCH_CHECK();
}
void fill_i64(long long *p, long long v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<long long><<<fill_grid(n), 256, 0, s>>>(p, v, n);
  RMX_LAUNCH_CHECK();
}
void iota_u32(uint32_t *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_iota<<<fill_grid(n), 256, 0, s>>>(p, n);
  RMX_LAUNCH_CHECK();
}

}  // namespace rmx
