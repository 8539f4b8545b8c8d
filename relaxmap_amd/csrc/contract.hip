// Module contraction (Network::convertModulesToSuperNodes, Module.cpp:3008-3077,
// SuperNode(Module, int), Node.cpp:76-106) and member bookkeeping
// (Network::updateMembersInModule, Module.cpp:2881-2902).
//
// Contraction always works on the original arcs and the nodes' current
// modules, as the reference does.  Relabel: super node id = rank of the
// module among non-empty modules (exclusive scan of non-empty flags).  Arcs
// between different super nodes get the key (src_sp, dst_sp); a stable LSD
// radix sort by key keeps, inside every key, the original arc order
// (source node ascending, then stored out-link order), which is the order in
// which the reference's std::map sums them.  Each key segment is then summed
// left to right by one thread, so the FP64 super-arc weights are bit-exact;
// the fixed-point sums the sweeps use are carried beside them.  In-lists are
// the transpose, again by a stable sort (ascending source).
#include "devutil.h"

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

// flag arcs whose ends lie in different super nodes
template <int G>
__global__ void k_inter_flags(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx, int32_t n,
                              const uint32_t *__restrict__ node_sp, uint32_t *__restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n) return;
  const uint32_t sa = node_sp[a];
  for (uint32_t k = optr[a] + lane; k < optr[a + 1]; k += G) flag[k] = node_sp[oidx[k]] != sa ? 1u : 0u;
}

template <int G>
__global__ void k_inter_keys(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx, int32_t n,
                             const uint32_t *__restrict__ node_sp, const uint32_t *__restrict__ flag,
                             const uint32_t *__restrict__ pos, int bits, uint64_t *__restrict__ keys,
                             uint32_t *__restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n) return;
  const uint64_t sa = node_sp[a];
  for (uint32_t k = optr[a] + lane; k < optr[a + 1]; k += G) {
    if (!flag[k]) continue;
    const uint32_t j = pos[k];
    keys[j] = (sa << bits) | (uint64_t)node_sp[oidx[k]];
    vals[j] = k;
  }
}

__global__ void k_key_heads(const uint64_t *__restrict__ keys, int64_t m, uint32_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_key_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, int64_t m,
                             uint32_t *__restrict__ start) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flag[i]) start[pos[i]] = (uint32_t)i;
}

// one thread per (src, dst) super arc: sequential FP64 sum in arc order
__global__ void k_reduce_segments(const uint32_t *__restrict__ start, int64_t n_seg, int64_t m,
                                  const uint64_t *__restrict__ keys, const uint32_t *__restrict__ arc,
                                  const double *__restrict__ w, const float *__restrict__ wf, int mixed, int bits,
                                  uint32_t *__restrict__ out_idx, double *__restrict__ out_w,
                                  long long *__restrict__ out_wx, uint32_t *__restrict__ src_cnt) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const uint32_t b = start[s];
  const uint32_t e = (s + 1 < n_seg) ? start[s + 1] : (uint32_t)m;
  double sum = 0.0;
  long long sx = 0;
  for (uint32_t i = b; i < e; ++i) {
    const uint32_t k = arc[i];
    const double f = w[k];
    sum += f;
    sx += fx_of(mixed ? (double)wf[k] : f);
  }
  const uint64_t key = keys[b];
  out_idx[s] = (uint32_t)(key & ((1ull << bits) - 1ull));
  out_w[s] = sum;
  out_wx[s] = sx;
  atomicAdd(&src_cnt[key >> bits], 1u);
}

__global__ void k_scatter_in(const uint32_t *__restrict__ in_of_out, int64_t m, const double *__restrict__ ow,
                             const long long *__restrict__ owx, const float *__restrict__ owf, double *__restrict__ iw,
                             long long *__restrict__ iwx, float *__restrict__ iwf) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const uint32_t k = in_of_out[e];
  iw[k] = ow[e];
  iwx[k] = owx[e];
  iwf[k] = owf[e];
}

__global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_propagate(const uint32_t *__restrict__ node_sp, const uint32_t *__restrict__ smod, int32_t n,
                            uint32_t *__restrict__ mod) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mod[i] = smod[node_sp[i]];
}

__global__ void k_copy_i32(const int32_t *__restrict__ a, int32_t *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int bits_for(int64_t n) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

}  // namespace

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
  c.level = 0;
  Level &L0 = c.L0;
  Level &L1 = c.L1;
  const int32_t n = c.N;
  const int64_t m = L0.m;

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
  L1.n = (int32_t)ns;   // level arrays keep their blocks across contractions
  L1.us.alloc(ns);
  L1.nmem.alloc(ns);
  L1.mod.alloc(ns);
  k_super_stats<<<grid_for(n, 256), 256, 0, s>>>(flag.p, mod2sp.p, n, c.ms.p, c.mNodes.p, L1.us.p, L1.nmem.p,
                                                 L1.mod.p, c.mUnits.p);
  RMX_LAUNCH_CHECK();
  c.node_sp.alloc(n);
  k_node_sp<<<grid_for(n, 256), 256, 0, s>>>(L0.mod.p, mod2sp.p, n, c.node_sp.p);
  RMX_LAUNCH_CHECK();
  c.sync();
  mod2sp.free();

  // inter-super-node arcs in original arc order
  DevBuf<uint32_t> aflag(c.cache), apos(c.cache), atot(c.cache);
  aflag.alloc(m);
  apos.alloc(m);
  atot.alloc(1);
  uint32_t m_inter = 0;
  if (m > 0) {
    k_inter_flags<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L0.out_ptr.p, L0.out_idx.p, n, c.node_sp.p,
                                                                      aflag.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(aflag.p, apos.p, m, atot.p, c.scratch, s);
    RMX_HIP(hipMemcpyAsync(&m_inter, atot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    c.sync();
  }
  const int bits = bits_for(ns);
  DevBuf<uint64_t> keys(c.cache);
  DevBuf<uint32_t> arcs(c.cache);
  keys.alloc(m_inter);
  arcs.alloc(m_inter);
  if (m_inter > 0) {
    k_inter_keys<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L0.out_ptr.p, L0.out_idx.p, n, c.node_sp.p,
                                                                     aflag.p, apos.p, bits, keys.p, arcs.p);
    RMX_LAUNCH_CHECK();
    radix_sort_pairs_u64(keys.p, arcs.p, m_inter, 0, 2 * bits, c.scratch, s);
  }
  c.sync();
  aflag.free();
  apos.free();

  // one super arc per distinct key
  uint32_t m_super = 0;
  DevBuf<uint32_t> kflag(c.cache), kpos(c.cache), kstart(c.cache), src_cnt(c.cache);
  src_cnt.alloc((int64_t)ns + 1);
  fill_u32(src_cnt.p, 0u, (int64_t)ns + 1, s);
  if (m_inter > 0) {
    kflag.alloc(m_inter);
    kpos.alloc(m_inter);
    k_key_heads<<<grid_for(m_inter, 256), 256, 0, s>>>(keys.p, m_inter, kflag.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(kflag.p, kpos.p, m_inter, atot.p, c.scratch, s);
    RMX_HIP(hipMemcpyAsync(&m_super, atot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    c.sync();
    kstart.alloc(m_super);
    k_key_starts<<<grid_for(m_inter, 256), 256, 0, s>>>(kflag.p, kpos.p, m_inter, kstart.p);
    RMX_LAUNCH_CHECK();
    c.sync();
    kflag.free();
    kpos.free();
  }
  L1.m = m_super;
  L1.out_idx.alloc(m_super);
  L1.out_w.alloc(m_super);
  L1.out_wx.alloc(m_super);
  L1.out_wf.alloc(m_super);
  if (m_super > 0) {
    k_reduce_segments<<<grid_for(m_super, 256), 256, 0, s>>>(kstart.p, m_super, m_inter, keys.p, arcs.p, L0.out_w.p,
                                                              L0.out_wf.p, c.precision == 1 ? 1 : 0, bits,
                                                              L1.out_idx.p, L1.out_w.p, L1.out_wx.p, src_cnt.p);
    RMX_LAUNCH_CHECK();
    k_to_float<<<grid_for(m_super, 256), 256, 0, s>>>(L1.out_w.p, L1.out_wf.p, m_super);
    RMX_LAUNCH_CHECK();
  }
  c.sync();
  keys.free();
  arcs.free();
  kstart.free();
  L1.out_ptr.alloc((int64_t)ns + 1);
  exclusive_scan_u32(src_cnt.p, L1.out_ptr.p, (int64_t)ns + 1, nullptr, c.scratch, s);
  c.sync();

  // transpose
  DevBuf<uint32_t> in_of_out(c.cache);
  build_in_csr(c, L1, &in_of_out);
  L1.in_w.alloc(m_super);
  L1.in_wx.alloc(m_super);
  L1.in_wf.alloc(m_super);
  if (m_super > 0) {
    k_scatter_in<<<grid_for(m_super, 256), 256, 0, s>>>(in_of_out.p, m_super, L1.out_w.p, L1.out_wx.p, L1.out_wf.p,
                                                         L1.in_w.p, L1.in_wx.p, L1.in_wf.p);
    RMX_LAUNCH_CHECK();
  }
  build_buckets(c, L1);
  c.level = 1;
  c.sync();
}

}  // namespace rmx
