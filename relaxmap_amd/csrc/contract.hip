// Module contraction (Network::convertModulesToSuperNodes, Module.cpp:3008-3077,
// SuperNode(Module, int), Node.cpp:76-106) and member bookkeeping
// (Network::updateMembersInModule, Module.cpp:2881-2902).
//
// Relabel: super node id = rank of the module among non-empty modules
// (exclusive scan of non-empty flags).  Every arc gets the key
// (src_sp, dst_sp), or a sentinel above every valid key when both ends lie in
// one super node; a stable LSD radix sort by key keeps, inside every key, the
// source level's arc order, and moves the sentinels to the end.  Each key
// segment is then summed left to right by one thread; the fixed-point sums
// the sweeps use are carried beside the FP64 sums.  In-lists are the
// transpose: the super arcs stably sorted by head (ascending source within a
// head), with the fixed-point weights carried through the sort.
//
// The reference always contracts the original arcs (source node ascending,
// then stored out-link order: the order its std::map sums them), so from
// the node level the FP64 super-arc weights are bit-exact.  When the current
// level is already a contraction, the new one is built from the current
// super arcs instead (a fraction of the original arcs): the fixed-point sums
// are exact either way, and the FP64 weights differ from summing the
// original arcs by rounding only.
#include "devutil.h"

#include <algorithm>
#include <vector>

namespace rmx {

namespace {

__global__ void k_nonempty(const int32_t *__restrict__ mNodes, int32_t n, uint32_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = mNodes[i] > 0 ? 1u : 0u;
}

__global__ void k_super_stats(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ mod2sp, int32_t n,
                              const Stat4 *__restrict__ ms, const int32_t *__restrict__ mNodes,
                              Stat4 *__restrict__ us, int32_t *__restrict__ nmem, uint32_t *__restrict__ smod, int32_t *__restrict__ mUnits) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  if (!flag[m]) {
    mUnits[m] = 0;
    return;
  }
  const uint32_t s = mod2sp[m];
  us[s] = ms[m];
  nmem[s] = mNodes[m];
  smod[s] = (uint32_t)m;
  mUnits[m] = 1;
}

__global__ void k_node_sp(const uint32_t *__restrict__ mod, const uint32_t *__restrict__ mod2sp, int32_t n,
                          uint32_t *__restrict__ node_sp) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) node_sp[i] = mod2sp[mod[i]];
}

// sort key of every arc: (src_sp, dst_sp), or kIntra when both ends lie in
// one super node (bits leaves the all-ones pattern unused by valid keys)
constexpr uint64_t kIntra = ~0ull;

// The value sorted with each key: the arc's index, or (carry_wf) the bits
// of its FP32 flow, which is then all the reduce needs of it.
template <int G>
__global__ void k_arc_keys(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx, int32_t n,
                           const uint32_t *__restrict__ node_sp, int bits, const float *__restrict__ carry_wf,
                           uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n || optr[a + 1] - optr[a] > kBigRow) return;   // big rows: k_arc_keys_big
  const uint32_t sa = node_sp[a];
  for (uint32_t k = optr[a] + lane; k < optr[a + 1]; k += G) {
    const uint32_t sb = node_sp[oidx[k]];
    keys[k] = sb != sa ? (((uint64_t)sa << bits) | (uint64_t)sb) : kIntra;
    vals[k] = carry_wf ? __float_as_uint(carry_wf[k]) : k;
  }
}

// chunks of the rows of more than kBigRow arcs (the order of the chunks is
// that of the atomic counter; every chunk's work is independent)
__global__ void k_big_rows(const uint32_t *__restrict__ ptr, int32_t n, BigRows B) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t b = ptr[r], d = ptr[r + 1] - b;
  if (d <= kBigRow) return;
  const uint32_t nch = (d + kBigRow - 1) / kBigRow;
  const uint32_t at = atomicAdd(B.ctr, nch);
  atomicAdd(B.arcs, (unsigned long long)d);
  for (uint32_t k = 0; k < nch; ++k) {
    B.crow[at + k] = (uint32_t)r;
    B.cbeg[at + k] = b + k * kBigRow;
  }
}

// the arcs of rows of more than kBigRow arcs, left out by k_arc_keys
__global__ __launch_bounds__(256) void k_arc_keys_big(BigRows B, const uint32_t *__restrict__ optr,
                                                      const uint32_t *__restrict__ oidx,
                                                      const uint32_t *__restrict__ node_sp, int bits,
                                                      const float *__restrict__ carry_wf,
                                                      uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t nch = *B.ctr;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const uint32_t a = B.crow[c], b = B.cbeg[c];
    const uint32_t e = b + kBigRow < optr[a + 1] ? b + kBigRow : optr[a + 1];
    const uint32_t sa = node_sp[a];
    for (uint32_t k = b + threadIdx.x; k < e; k += 256) {
      const uint32_t sb = node_sp[oidx[k]];
      keys[k] = sb != sa ? (((uint64_t)sa << bits) | (uint64_t)sb) : kIntra;
      vals[k] = carry_wf ? __float_as_uint(carry_wf[k]) : k;
    }
  }
}

// number of keys before the first kIntra (keys sorted ascending)
__global__ void k_count_inter(const uint64_t *__restrict__ keys, int64_t m, uint32_t *__restrict__ out) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] == kIntra) hi = mid;
    else lo = mid + 1;
  }
  *out = (uint32_t)lo;
}

// Segments of the sorted inter keys (one per super arc), in two reads of the
// keys: k_arc_seg_count counts each tile's heads (positions whose key differs
// from the one before), and after a scan of the tile counts k_arc_seg_starts
// writes every segment's first position and, where the source super node
// changes between two segments, the out-pointers of the sources in between.
// A tile is 16 rounds of 256 consecutive keys; a head's rank inside its tile
// follows (round, wavefront, lane) order, which is position order.
constexpr int kSegItems = 16;
constexpr int kSegTile = 256 * kSegItems;

__device__ inline bool arc_seg_head(const uint64_t *__restrict__ keys, int64_t j, uint64_t &k, uint64_t &kp) {
  k = keys[j];
  kp = j > 0 ? keys[j - 1] : ~k;
  return k != kp;
}

__global__ __launch_bounds__(256) void k_arc_seg_count(const uint64_t *__restrict__ keys, int64_t m,
                                                   uint32_t *__restrict__ tile_cnt) {
  __shared__ uint32_t lds[256 / kWave];
  const int64_t base = (int64_t)blockIdx.x * kSegTile;
  uint32_t c = 0;
#pragma unroll 4
  for (int i = 0; i < kSegItems; ++i) {
    const int64_t j = base + (int64_t)i * 256 + threadIdx.x;
    uint64_t k, kp;
    if (j < m && arc_seg_head(keys, j, k, kp)) ++c;
  }
  c = block_sum<uint32_t, 256>(c, lds);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = c;
}

// tile_base: exclusive scan of the tile counts; *n_seg: their total
__global__ __launch_bounds__(256) void k_arc_seg_starts(const uint64_t *__restrict__ keys, int64_t m, int bits,
                                                    int32_t ns, const uint32_t *__restrict__ tile_base,
                                                    const uint32_t *__restrict__ n_seg, uint32_t *__restrict__ start,
                                                    uint32_t *__restrict__ out_ptr) {
  constexpr int NW = 256 / kWave;
  static_assert(kSegItems * NW == kWave, "one lane per (round, wavefront) count");
  __shared__ uint32_t cnt[kSegItems * NW];   // heads of (round, wavefront), then their exclusive prefix
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const int64_t base = (int64_t)blockIdx.x * kSegTile;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  uint64_t key[kSegItems], prev[kSegItems];
  uint32_t rank[kSegItems];
  uint32_t heads = 0;   // bit i: this lane's element of round i is a head
#pragma unroll
  for (int i = 0; i < kSegItems; ++i) {
    const int64_t j = base + (int64_t)i * 256 + t;
    const bool h = j < m && arc_seg_head(keys, j, key[i], prev[i]);
    const uint64_t bal = __ballot(h);
    if (lane == 0) cnt[i * NW + w] = (uint32_t)__popcll(bal);
    rank[i] = (uint32_t)__popcll(bal & below);
    heads |= (h ? 1u : 0u) << i;
  }
  __syncthreads();
  if (w == 0) {   // exclusive scan of the 64 (round, wavefront) counts, one per lane
    const uint32_t v = cnt[lane];
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, kWave);
      if (lane >= off) x += y;
    }
    cnt[lane] = x - v;
  }
  __syncthreads();
  const uint32_t tb = tile_base[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kSegItems; ++i) {
    if (!((heads >> i) & 1u)) continue;
    const int64_t j = base + (int64_t)i * 256 + t;
    const uint32_t s = tb + cnt[i * NW + w] + rank[i];
    start[s] = (uint32_t)j;
    const int64_t src = (int64_t)(key[i] >> bits);
    const int64_t src_prev = j > 0 ? (int64_t)(prev[i] >> bits) : -1;
    for (int64_t d = src_prev + 1; d <= src; ++d) out_ptr[d] = s;   // empty when the source is unchanged
  }
  if (m > 0 && base <= m - 1 && m - 1 < base + kSegTile && (m - 1 - base) % 256 == t) {
    // the last key: super nodes after its source have no out-arcs
    const uint32_t total = *n_seg;
    for (int64_t d = (int64_t)(keys[m - 1] >> bits) + 1; d <= ns; ++d) out_ptr[d] = total;
  }
}

// one thread per (src, dst) super arc: sequential FP64 sum in arc order (in
// mixed precision, from the node level: of the FP32 arc flows, carried by
// the sort in place of the arc index); also the transpose's key (dst, src)
// and the fixed-point weight it carries
__global__ void k_reduce_segments(const uint32_t *__restrict__ start, int64_t n_seg, int64_t m,
                                  const uint64_t *__restrict__ keys, const uint32_t *__restrict__ arc,
                                  const double *__restrict__ w, const long long *__restrict__ wx, int mixed, int bits,
                                  uint32_t *__restrict__ out_idx, double *__restrict__ out_w,
                                  long long *__restrict__ out_wx, uint64_t *__restrict__ tkey,
                                  uint64_t *__restrict__ twx) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const uint32_t b = start[s];
  const uint32_t e = (s + 1 < n_seg) ? start[s + 1] : (uint32_t)m;
  double sum = 0.0;
  long long sx = 0;
  // kRowAhead elements in flight, added in order: a long segment (two large
  // super nodes joined by many arcs) does not cost one round trip per arc
  for (uint32_t i0 = b; i0 < e; i0 += kRowAhead) {
    uint32_t k[kRowAhead];
    double f[kRowAhead];
    long long fx[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) k[j] = i0 + j < e ? arc[i0 + j] : 0u;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) {
      // mixed precision reads only the FP32 flow of a node-level arc: its
      // FP64 sum and its fixed-point value (the one the sweeps use) both
      // come from it
      f[j] = i0 + j < e ? ((mixed && !wx) ? (double)__uint_as_float(k[j]) : w[k[j]]) : 0.0;
      fx[j] = (wx && i0 + j < e) ? wx[k[j]] : 0ll;
    }
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) {
      if (i0 + j >= e) break;
      sum += f[j];
      sx += wx ? fx[j] : fx_of(f[j]);
    }
  }
  const uint64_t key = keys[b];
  const uint64_t mask = (1ull << bits) - 1ull;
  out_idx[s] = (uint32_t)(key & mask);
  out_w[s] = sum;
  out_wx[s] = sx;
  tkey[s] = ((key & mask) << bits) | (key >> bits);
  twx[s] = (uint64_t)sx;
}

// in-CSR of the new level from the transposed keys (head, source) sorted by
// head: sources and fixed-point weights
__global__ void k_in_from_keys(const uint64_t *__restrict__ key, const uint64_t *__restrict__ wx, int64_t m,
                               int bits, uint32_t *__restrict__ in_idx, long long *__restrict__ in_wx) {
  const uint64_t mask = (1ull << bits) - 1ull;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    in_idx[k] = (uint32_t)(key[k] & mask);
    in_wx[k] = (long long)wx[k];
  }
}

// in_ptr from the heads (key >> bits, ascending): position i fills every
// head in (head[i-1], head[i]], and position m the tail up to n
__global__ void k_in_ptr(const uint64_t *__restrict__ key, int bits, int64_t m, int32_t n,
                         uint32_t *__restrict__ ptr) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t lo = i == 0 ? 0 : (int64_t)(key[i - 1] >> bits) + 1;
    const int64_t hi = i == m ? (int64_t)n : (int64_t)(key[i] >> bits);
    for (int64_t d = lo; d <= hi; ++d) ptr[d] = (uint32_t)i;
  }
}

__global__ void k_propagate(const uint32_t *__restrict__ node_sp, const uint32_t *__restrict__ smod, int32_t n,
                            uint32_t *__restrict__ mod) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mod[i] = smod[node_sp[i]];
}

// node -> new super node through the previous super node
__global__ void k_compose(uint32_t *__restrict__ node_sp, const uint32_t *__restrict__ usp, int32_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) node_sp[i] = usp[node_sp[i]];
}

__global__ void k_copy_i32(const int32_t *__restrict__ a, int32_t *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// a batch: the instance of every super node (that of its module id) and the
// first super node of every instance (super nodes follow module-id order,
// and an instance's modules lie in its id range)
__global__ void k_unit_inst(const uint32_t *__restrict__ mod, int32_t ns, const uint32_t *__restrict__ b_node,
                            uint32_t *__restrict__ uinst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) uinst[i] = b_node[mod[i]];
}

__global__ void k_unit_off(const uint32_t *__restrict__ uinst, int32_t ns, int32_t nb, uint32_t *__restrict__ uoff) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I > nb) return;
  uint32_t lo = 0, hi = (uint32_t)ns;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (uinst[mid] < (uint32_t)I) lo = mid + 1;
    else hi = mid;
  }
  uoff[I] = lo;
}

int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

}  // namespace

void plan_big_rows(const uint32_t *ptr, int32_t n, const BigRows &B, hipStream_t s) {
  fill_u32(B.ctr, 0u, 1, s);
  fill_i64((long long *)B.arcs, 0, 1, s);
  if (n <= 0) return;
  k_big_rows<<<grid_for(n, 256), 256, 0, s>>>(ptr, n, B);
  RMX_LAUNCH_CHECK();
}

void update_members(Ctx &c) {
  k_copy_i32<<<grid_for(c.N, 256), 256, 0, c.stream>>>(c.mNodes.p, c.mStale.p, c.N);
  RMX_LAUNCH_CHECK();
}

void propagate_labels(Ctx &c) {
  if (c.level == 0) return;
  k_propagate<<<grid_for(c.N, 256), 256, 0, c.stream>>>(c.node_sp.p, c.L1.mod.p, c.N, c.L0.mod.p);
  RMX_LAUNCH_CHECK();
}

void contract(Ctx &c) {
  hipStream_t s = c.stream;
  propagate_labels(c);
  const bool from_super = c.level == 1;
  c.level = 0;
  Level &L0 = c.L0;
  Level &S = from_super ? c.L1 : c.L0;   // level whose arcs are contracted
  Level &D = c.L2;                       // the new super level, swapped into L1 at the end
  const int32_t n = c.N;
  const int32_t ns_src = S.n;
  const int64_t m = S.m;

  // relabel non-empty modules
  DevBuf<uint32_t> flag(c.cache), mod2sp(c.cache), total(c.cache);
  flag.alloc(n);
  mod2sp.alloc(n);
  total.alloc(1);
  k_nonempty<<<grid_for(n, 256), 256, 0, s>>>(c.mNodes.p, n, flag.p);
  RMX_LAUNCH_CHECK();
  exclusive_scan_u32(flag.p, mod2sp.p, n, total.p, c.scratch, s);
  uint32_t ns = 0;
  RMX_HIP(hipMemcpyAsync(&ns, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  c.sync();
  D.n = (int32_t)ns;   // level arrays keep their blocks across contractions
  D.us.alloc(ns);
  D.nmem.alloc(ns);
  D.mod.alloc(ns);
  k_super_stats<<<grid_for(n, 256), 256, 0, s>>>(flag.p, mod2sp.p, n, c.ms.p, c.mNodes.p, D.us.p, D.nmem.p,
                                                 D.mod.p, c.mUnits.p);
  RMX_LAUNCH_CHECK();
  // unit of the source level -> new super node
  DevBuf<uint32_t> usp_buf(c.cache);
  const uint32_t *usp;
  if (from_super) {
    usp_buf.alloc(ns_src);
    k_node_sp<<<grid_for(ns_src, 256), 256, 0, s>>>(S.mod.p, mod2sp.p, ns_src, usp_buf.p);
    RMX_LAUNCH_CHECK();
    k_compose<<<grid_for(n, 256), 256, 0, s>>>(c.node_sp.p, usp_buf.p, n);
    RMX_LAUNCH_CHECK();
    usp = usp_buf.p;
  } else {
    c.node_sp.alloc(n);
    k_node_sp<<<grid_for(n, 256), 256, 0, s>>>(L0.mod.p, mod2sp.p, n, c.node_sp.p);
    RMX_LAUNCH_CHECK();
    usp = c.node_sp.p;
  }
  c.sync();
  mod2sp.free();

  // every arc by key: the inter-super-node arcs in key order (source arc
  // order inside a key), the intra ones after them
  DevBuf<uint32_t> atot(c.cache);
  atot.alloc(1);
  const int bits = bits_for((int64_t)ns + 1);   // the all-ones key stays free for kIntra
  DevBuf<uint64_t> keys(c.cache);
  DevBuf<uint32_t> arcs(c.cache);
  keys.alloc(m);
  arcs.alloc(m);
  uint32_t m_inter = 0;
  if (m > 0) {
    const bool carry = c.precision == 1 && !from_super;   // node level, mixed: the sort carries the FP32 flows
    DevBuf<uint32_t> bctr(c.cache), bcrow(c.cache), bcbeg(c.cache);
    DevBuf<long long> barcs(c.cache);
    bctr.alloc(1);
    barcs.alloc(1);
    bcrow.alloc(big_rows_cap(m));
    bcbeg.alloc(big_rows_cap(m));
    const BigRows big{bctr.p, bcrow.p, bcbeg.p, (unsigned long long *)barcs.p};
    plan_big_rows(S.out_ptr.p, ns_src, big, s);
    long long m_big = 0;   // arcs of the big rows (profiling mode)
    if (c.prof.on) {
      RMX_HIP(hipMemcpyAsync(&m_big, barcs.p, sizeof(long long), hipMemcpyDeviceToHost, s));
      c.sync();
    }
    // per arc: head id, its super node (gathered), key and value written,
    // the FP32 flow when carried; per unit: bounds and its super node
    const double per_arc = 4 + 4 + 8 + 4 + (carry ? 4 : 0);
    RMX_PROF_BEGIN(c);
    k_arc_keys<16><<<grid_for((int64_t)ns_src * 16, 256), 256, 0, s>>>(S.out_ptr.p, S.out_idx.p, ns_src, usp, bits,
                                                                        carry ? S.out_wf.p : nullptr, keys.p, arcs.p);
    RMX_LAUNCH_CHECK();
    RMX_PROF_END(c, "k_arc_keys<16>", (double)(m - m_big) * per_arc + (double)ns_src * 12);
    RMX_PROF_BEGIN(c);
    k_arc_keys_big<<<big_rows_grid(m), 256, 0, s>>>(big, S.out_ptr.p, S.out_idx.p, usp, bits,
                                                     carry ? S.out_wf.p : nullptr, keys.p, arcs.p);
    RMX_LAUNCH_CHECK();
    RMX_PROF_END(c, "k_arc_keys_big", (double)m_big * per_arc);
    radix_sort_pairs_u64(keys.p, arcs.p, m, 0, 2 * bits, c.scratch, s, c.prof.on ? &c.prof : nullptr);
    k_count_inter<<<1, 1, 0, s>>>(keys.p, m, atot.p);
    RMX_LAUNCH_CHECK();
    RMX_HIP(hipMemcpyAsync(&m_inter, atot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  c.sync();

  // one super arc per distinct key: its first position, and the out-pointers
  uint32_t m_super = 0;
  DevBuf<uint32_t> tcnt(c.cache), kstart(c.cache);
  D.out_ptr.alloc((int64_t)ns + 1);
  if (m_inter > 0) {
    const int64_t nt = ((int64_t)m_inter + kSegTile - 1) / kSegTile;
    tcnt.alloc(nt);
    kstart.alloc(m_inter);   // capacity: at most one segment per key
    RMX_PROF_BEGIN(c);
    k_arc_seg_count<<<(unsigned)nt, 256, 0, s>>>(keys.p, m_inter, tcnt.p);
    RMX_LAUNCH_CHECK();
    RMX_PROF_END(c, "k_arc_seg_count", (double)m_inter * 8 + (double)nt * 4);
    exclusive_scan_u32(tcnt.p, tcnt.p, nt, atot.p, c.scratch, s);
    RMX_PROF_BEGIN(c);
    k_arc_seg_starts<<<(unsigned)nt, 256, 0, s>>>(keys.p, m_inter, bits, (int32_t)ns, tcnt.p, atot.p, kstart.p,
                                               D.out_ptr.p);
    RMX_LAUNCH_CHECK();
    const float ms_starts = c.prof.on ? c.prof.stop(s) : 0.f;
    RMX_HIP(hipMemcpyAsync(&m_super, atot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    c.sync();
    // per key: the key read; per segment: its start written; per super node: its out-pointer
    if (c.prof.on) c.prof.add("k_arc_seg_starts", ms_starts, (double)m_inter * 8 + (double)m_super * 4 + (double)ns * 4);
    tcnt.free();
  } else {
    fill_u32(D.out_ptr.p, 0u, (int64_t)ns + 1, s);
  }
  D.m = m_super;
  D.out_idx.alloc(m_super);
  D.out_w.alloc(m_super);
  D.out_wx.alloc(m_super);
  DevBuf<uint64_t> tkey(c.cache), twx(c.cache);   // transposed keys and their weights
  tkey.alloc(m_super);
  twx.alloc(m_super);

  if (m_super > 0) {
    RMX_PROF_BEGIN(c);
    k_reduce_segments<<<grid_for(m_super, 256), 256, 0, s>>>(
        kstart.p, m_super, m_inter, keys.p, arcs.p, S.out_w.p, from_super ? S.out_wx.p : nullptr,
        c.precision == 1 ? 1 : 0, bits, D.out_idx.p, D.out_w.p, D.out_wx.p, tkey.p, twx.p);
    RMX_LAUNCH_CHECK();
    // per inter arc: its value and the weights it names (none when the sort
    // carried the FP32 flow, FP64 + fixed point from a super level, FP64
    // from the node level); per super arc: two starts, its key, five outputs
    const double per_arc = 4.0 + ((c.precision == 1 && !from_super) ? 0.0 : (from_super ? 16.0 : 8.0));
    RMX_PROF_END(c, "k_reduce_segments", (double)m_inter * per_arc + (double)m_super * (8 + 8 + 4 + 8 + 8 + 8 + 8));
  }
  c.sync();
  keys.free();
  arcs.free();
  kstart.free();
  usp_buf.free();

  // transpose: stable sort by head (the high half of the transposed key)
  D.in_ptr.alloc((int64_t)ns + 1);
  D.in_idx.alloc(m_super);
  D.in_wx.alloc(m_super);
  if (m_super > 0) {
    uint64_t *sk = nullptr, *sw = nullptr;
    radix_sort_pairs_u64_u64(tkey.p, twx.p, m_super, bits, 2 * bits, c.scratch, s, &sk, &sw,
                             c.prof.on ? &c.prof : nullptr);
    RMX_PROF_BEGIN(c);
    k_in_from_keys<<<(unsigned)std::min<int64_t>(16384, (m_super + 255) / 256), 256, 0, s>>>(sk, sw, m_super, bits,
                                                                                            D.in_idx.p, D.in_wx.p);
    RMX_LAUNCH_CHECK();
    RMX_PROF_END(c, "k_in_from_keys", (double)m_super * (8 + 8 + 4 + 8));
    k_in_ptr<<<(unsigned)std::min<int64_t>(16384, (m_super + 256) / 256), 256, 0, s>>>(sk, bits, m_super, (int32_t)ns,
                                                                                      D.in_ptr.p);
    RMX_LAUNCH_CHECK();
  } else {
    fill_u32(D.in_ptr.p, 0u, (int64_t)ns + 1, s);
  }
  c.sync();
  tkey.free();
  twx.free();
  if (c.nb > 0) {
    D.uinst.alloc(ns);
    D.uoff.alloc((int64_t)c.nb + 1);
    k_unit_inst<<<grid_for(ns, 256), 256, 0, s>>>(D.mod.p, (int32_t)ns, c.b_node.p, D.uinst.p);
    RMX_LAUNCH_CHECK();
    k_unit_off<<<grid_for((int64_t)c.nb + 1, 256), 256, 0, s>>>(D.uinst.p, (int32_t)ns, c.nb, D.uoff.p);
    RMX_LAUNCH_CHECK();
  }
  build_buckets(c, D);
  c.sync();
  c.L1.swap(c.L2);
  c.level = 1;
}

}  // namespace rmx
