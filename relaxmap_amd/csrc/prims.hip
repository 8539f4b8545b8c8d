// Scans and a stable LSD radix sort written for CDNA4 wave64.
//
// Scan: reduce-then-scan over tiles of 4096 elements (256 threads x 16),
// recursing on the tile sums; no inter-block spinning.
// Radix sort: stable LSD over 8-bit digits.  One kernel (k_sort_hist) reads
// the keys once and counts the digits of every pass; a 256-lane scan per pass
// (k_sort_digit_scan) turns them into each digit's first output position.
// Each pass is then one kernel (k_sort_onesweep): tiles are numbered in the
// order their workgroups start (atomic ticket), each wave ranks its share of
// the tile with ballots and a wave-private digit count, the tile publishes
// its digit counts and reorders itself by digit in LDS, and then finds its
// digits' offsets by a decoupled look-back over its predecessors' published
// counts (bounded waits that raise an error flag instead of hanging), and
// writes coalesced digit runs from LDS.
#include "prims.h"
#include "context.h"

namespace rmx {

void *Scratch::get(size_t need) {
  if (need > bytes) {
    if (fixed) throw HipError("fixed scratch arena too small");
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
  if (fixed) return;
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
constexpr int kRadix = 256;

// Elements per thread of a sort tile: 24 for pairs of up to 12 bytes (72 KB
// of LDS at 12 bytes, 2 workgroups per CU), 16 for 16-byte pairs (64 KB).
// On 500M pairs: 64-bit keys with 32-bit values 24.1 ms (16: 29.4, 20: 26.0),
// 64-bit values 16.8 ms (20: 20.7, 24: 18.9).
template <typename K, typename V>
struct SortShape {
  static constexpr int items = sizeof(K) + sizeof(V) >= 16 ? 16 : 24;
  static constexpr int tile = kSortBlock * items;
};

constexpr int kMaxPasses = 8;

// Digit counts of every pass at once: one read of the keys.  Keys often
// arrive grouped by their high bits (arcs in source order), so neighbouring
// lanes of a wavefront often share a digit: each run of equal digits is
// found with one shuffle and one ballot and added by its first lane, instead
// of one LDS atomic per lane on the same slot.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const K *__restrict__ keys, int64_t n, int begin_bit,
                                                          int passes, uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[kMaxPasses][kRadix];
#pragma unroll
  for (int q = 0; q < kMaxPasses; ++q) h[q][threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  constexpr int kHistLoads = 8;   // keys per lane in flight
  for (int64_t j0 = (int64_t)blockIdx.x * kSortBlock * kHistLoads; j0 < n;
       j0 += (int64_t)gridDim.x * kSortBlock * kHistLoads) {
    K kk[kHistLoads];
#pragma unroll
    for (int u = 0; u < kHistLoads; ++u) {
      const int64_t j = j0 + u * kSortBlock + threadIdx.x;
      kk[u] = j < n ? keys[j] : K(0);
    }
#pragma unroll
    for (int u = 0; u < kHistLoads; ++u) {
      const int64_t j = j0 + u * kSortBlock + threadIdx.x;
      const bool valid = j < n;   // the valid lanes of a wavefront are a prefix
      const uint64_t live = __ballot(valid);
      if (!live) continue;
      const K k = kk[u];
      const int n_live = __popcll(live);
      for (int q = 0; q < passes; ++q) {
        // a run of lanes with one digit adds its length with one atomic (its
        // first lane's): keys grouped by their high bits repeat a digit
        const uint32_t d = (uint32_t)(k >> (begin_bit + 8 * q)) & 0xffu;
        const uint32_t dp = __shfl_up(d, 1, kWave);
        const bool head = valid && (lane == 0 || d != dp);
        const uint64_t heads = __ballot(head);
        if (head) {
          const uint64_t above = lane == kWave - 1 ? 0ull : heads >> (lane + 1);
          const int next = above ? lane + 1 + __builtin_ctzll(above) : n_live;
          atomicAdd(&h[q][d], (uint32_t)(next - lane));
        }
      }
    }
  }
  __syncthreads();
  for (int q = 0; q < passes; ++q)
    if (h[q][threadIdx.x]) atomicAdd(&counts[q * kRadix + threadIdx.x], h[q][threadIdx.x]);
}

// counts of pass blockIdx.x -> the first output position of every digit
__global__ __launch_bounds__(kRadix) void k_sort_digit_scan(uint32_t *__restrict__ counts) {
  __shared__ uint32_t wsum[kRadix / kWave];
  uint32_t *c = counts + blockIdx.x * kRadix;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const uint32_t v = c[t];
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int q = 0; q < w; ++q) pre += wsum[q];
  c[t] = pre + x - v;
}

// Look-back words: the state in the top two bits, a count below.
constexpr uint64_t kLbAgg = 1ull << 62;   // the tile's own digit count
constexpr uint64_t kLbInc = 2ull << 62;   // the count of this tile and every earlier one
constexpr uint64_t kLbCount = (1ull << 62) - 1ull;
constexpr uint32_t kLbSpinLimit = 1u << 20;

__device__ inline uint64_t lb_load(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One pass of the stable sort, in one sweep.  Tiles are numbered in the order
// workgroups start (an atomic counter), so every tile's predecessors are
// already running.  Wave w owns a quarter of the tile's elements, in index
// order, taken 64 at a time; it ranks each element among the
// wave's earlier elements of the same digit with ballots (equal-digit peers
// of a step) and a wave-private digit count in LDS, so no workgroup barrier
// is needed while ranking.  The tile then publishes its digit counts and is
// reordered by digit in LDS; thread d then walks back over the predecessors'
// words (decoupled look-back) until one holds an inclusive count, which gives
// the tile's first output position per digit, and each digit's run is
// written with consecutive lanes on consecutive addresses.
// A wait that exceeds kLbSpinLimit polls sets *err and the tile writes
// nothing (the caller reports the error).
template <typename K, typename V>
__global__ __launch_bounds__(kSortBlock) void k_sort_onesweep(const K *__restrict__ keys_in,
                                                              const V *__restrict__ vals_in,
                                                              K *__restrict__ keys_out, V *__restrict__ vals_out,
                                                              int64_t n, int shift,
                                                              const uint32_t *__restrict__ digit_base,
                                                              uint64_t *__restrict__ look,
                                                              uint32_t *__restrict__ tile_ctr,
                                                              uint32_t *__restrict__ err) {
  constexpr int kSortItems = SortShape<K, V>::items;
  constexpr int kSortTile = SortShape<K, V>::tile;
  constexpr int W = kSortBlock / kWave;
  constexpr int PW = kSortTile / W;   // elements per wave
  __shared__ K s_key[kSortTile];
  __shared__ V s_val[kSortTile];
  __shared__ uint32_t hist[W][kRadix];   // per-wave digit counts, then per-wave first positions
  __shared__ uint32_t tile_off[kRadix];
  __shared__ uint32_t digit_start[kRadix];
  __shared__ uint32_t wsum[W];
  __shared__ uint32_t s_tile;
  __shared__ int s_bad;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  if (t == 0) {
    s_tile = atomicAdd(tile_ctr, 1u);
    s_bad = 0;
  }
#pragma unroll
  for (int q = 0; q < W; ++q) hist[q][t] = 0;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kSortTile;
  const int64_t wbase = base + (int64_t)w * PW;
  K key[kSortItems];
  V val[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const int64_t j = wbase + (int64_t)i * kWave + lane;
    const bool valid = j < n;
    key[i] = valid ? keys_in[j] : K(0);
    val[i] = valid ? vals_in[j] : V(0);
  }
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
  {   // thread t holds digit t: its count and first position in the tile, then per wave
    uint32_t c[W];
    uint32_t tot = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
      c[q] = hist[q][t];
      tot += c[q];
    }
    uint64_t *mine = look + tile * kRadix + t;
    lb_store(mine, (tile == 0 ? kLbInc : kLbAgg) | (uint64_t)tot);
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
    // The tile is reordered by digit in LDS before the look-back: that
    // needs only the tile's own positions, and while it runs the
    // predecessors publish further, so the walk back is shorter (one pass
    // of 500M pairs: 3.81 -> 3.57 ms).
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
    // decoupled look-back for digit t
    uint64_t excl = 0;
    if (tile > 0) {
      int64_t j = tile - 1;
      uint32_t spins = 0;
      while (true) {
        const uint64_t v = lb_load(look + j * kRadix + t);
        const uint64_t state = v & ~kLbCount;
        if (state == 0) {
          if (++spins > kLbSpinLimit) {
            atomicOr(err, 1u);
            s_bad = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += v & kLbCount;
        if (state == kLbInc || j == 0) break;
        --j;
      }
      lb_store(mine, kLbInc | (excl + tot));
    }
    tile_off[t] = digit_base[t] + (uint32_t)excl;
  }
  __syncthreads();
  if (s_bad) return;
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
// back into keys/vals.  Waits for the stream at the end, to report a failed
// look-back (which leaves the output incomplete) as an exception.
// Scratch layout of one sort: temporary keys and values, look-back words,
// digit counts, tile counters and the error flag.
struct SortLayout {
  size_t off_v, off_l, off_c, total;
  int64_t n_tiles, n_small;
};

template <typename K, typename V>
SortLayout sort_layout(int64_t n, int passes) {
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  SortLayout l;
  l.n_tiles = (n + SortShape<K, V>::tile - 1) / SortShape<K, V>::tile;
  l.n_small = (int64_t)passes * kRadix + passes + 1;
  l.off_v = align(sizeof(K) * n);
  l.off_l = l.off_v + align(sizeof(V) * n);
  l.off_c = l.off_l + align(sizeof(uint64_t) * l.n_tiles * kRadix);
  l.total = l.off_c + align(sizeof(uint32_t) * l.n_small);
  return l;
}

template <typename T>
const char *type_name();
template <>
const char *type_name<uint32_t>() { return "unsigned int"; }
template <>
const char *type_name<uint64_t>() { return "unsigned long"; }

template <typename K, typename V>
void radix_sort_impl(K *keys, V *vals, int64_t n, int begin_bit, int end_bit, Scratch &scratch, hipStream_t s,
                     K **out_k = nullptr, V **out_v = nullptr, Profiler *prof = nullptr) {
  if (out_k) *out_k = keys;
  if (out_v) *out_v = vals;
  if (n <= 1 || begin_bit >= end_bit) return;
  if (n > (int64_t)0xffffffffLL) throw HipError("radix sort: n >= 2^32");
  const int passes = (end_bit - begin_bit + 7) / 8;
  if (passes > kMaxPasses) throw HipError("radix sort: more than 64 key bits");
  const SortLayout lay = sort_layout<K, V>(n, passes);
  const int64_t n_tiles = lay.n_tiles;
  const int64_t n_small = lay.n_small;   // digit counts, tile counters, error flag
  const size_t off_k = 0, off_v = lay.off_v, off_l = lay.off_l, off_c = lay.off_c;
  char *base = (char *)scratch.get(lay.total);
  K *tk = (K *)(base + off_k);
  V *tv = (V *)(base + off_v);
  uint64_t *look = (uint64_t *)(base + off_l);
  uint32_t *counts = (uint32_t *)(base + off_c);
  uint32_t *ctr = counts + passes * kRadix;
  uint32_t *err = ctr + passes;
  fill_u32(counts, 0u, n_small, s);
  if (prof) prof->begin(s);
  k_sort_hist<K><<<(unsigned)std::min<int64_t>(n_tiles, 2048), kSortBlock, 0, s>>>(keys, n, begin_bit, passes, counts);
  RMX_LAUNCH_CHECK();
  if (prof) prof->end(s, std::string("k_sort_hist<") + type_name<K>() + ">", (double)n * sizeof(K));
  k_sort_digit_scan<<<(unsigned)passes, kRadix, 0, s>>>(counts);
  RMX_LAUNCH_CHECK();
  K *src_k = keys, *dst_k = tk;
  V *src_v = vals, *dst_v = tv;
  for (int q = 0; q < passes; ++q) {
    fill_u32((uint32_t *)look, 0u, 2 * n_tiles * kRadix, s);
    if (prof) prof->begin(s);
    k_sort_onesweep<K, V><<<(unsigned)n_tiles, kSortBlock, 0, s>>>(src_k, src_v, dst_k, dst_v, n, begin_bit + 8 * q,
                                                                   counts + q * kRadix, look, ctr + q, err);
    RMX_LAUNCH_CHECK();
    if (prof)
      prof->end(s, std::string("k_sort_onesweep<") + type_name<K>() + ", " + type_name<V>() + ">",
                2.0 * (double)n * (sizeof(K) + sizeof(V)));
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
  uint32_t h_err = 0;
  RMX_HIP(hipMemcpyAsync(&h_err, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RMX_HIP(hipStreamSynchronize(s));
  if (h_err) throw HipError("radix sort: look-back wait exceeded its limit");
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

void exclusive_scan_reserve(int64_t n, Scratch &scratch) { scratch.get(sizeof(uint32_t) * scan_tmp_elems(n)); }

size_t exclusive_scan_scratch_bytes(int64_t n) { return sizeof(uint32_t) * scan_tmp_elems(n); }

size_t radix_sort_scratch_bytes(int64_t n, int key_bytes, int val_bytes, int bits) {
  if (n <= 1 || bits <= 0) return 0;
  const int passes = (bits + 7) / 8;
  if (key_bytes == 4 && val_bytes == 4) return sort_layout<uint32_t, uint32_t>(n, passes).total;
  if (key_bytes == 8 && val_bytes == 4) return sort_layout<uint64_t, uint32_t>(n, passes).total;
  if (key_bytes == 8 && val_bytes == 8) return sort_layout<uint64_t, uint64_t>(n, passes).total;
  throw HipError("radix sort: no such key / value width");
}

void radix_sort_pairs_u32(uint32_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s, Profiler *prof) {
  radix_sort_impl<uint32_t, uint32_t>(keys, vals, n, begin_bit, end_bit, scratch, s, nullptr, nullptr, prof);
}

void radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s, Profiler *prof) {
  radix_sort_impl<uint64_t, uint32_t>(keys, vals, n, begin_bit, end_bit, scratch, s, nullptr, nullptr, prof);
}

void radix_sort_pairs_u64_u64(uint64_t *keys, uint64_t *vals, int64_t n, int begin_bit, int end_bit,
                              Scratch &scratch, hipStream_t s, uint64_t **out_keys, uint64_t **out_vals,
                              Profiler *prof) {
  radix_sort_impl<uint64_t, uint64_t>(keys, vals, n, begin_bit, end_bit, scratch, s, out_keys, out_vals, prof);
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
