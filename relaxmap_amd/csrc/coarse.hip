// Coarse tune: the sub-module partition of every module and the super nodes
// built from the sub-modules.
//
// generate_sub_modules (OmpRelaxmap.cpp:500-632) cuts every module out of
// the network (generate_network_from_module, OmpRelaxmap.cpp:640-783: the
// members, the arcs between them, their flows, teleport weights and dangling
// flows unchanged), partitions each such sub-network from singletons with
// its own codelength (partition_module_network, OmpRelaxmap.cpp:436-498:
// sweeps until a sweep gains less than the threshold, at most maxIter), and
// makes every resulting module a super node that starts in its parent
// module (Network::generateSuperNodesFromSubModules, Module.cpp:3083-3158).
//
// Here all sub-networks are solved at once, as one batch of instances over
// the node arrays: the level of intra-module arcs replaces the node level,
// a second set of module statistics holds the sub-modules, and every
// instance has its own S (sum of its sub-module exits) and L, reduced by one
// wavefront per instance in a fixed order.  An instance stops when its own
// sweep gain falls below the threshold.  The sweeps are the node-level
// sweep kernels; no empty module is offered (the reference's stale member
// lists hold one member throughout, Module.cpp:1073-1077).
//
// The super nodes' exit is the consistent (alpha P + beta D)(1 - T) + beta X
// with X the flow on arcs leaving the sub-module, rather than the reference's
// per-member teleport sum (Module.cpp:3136-3148).
#include "devutil.h"

#include <cmath>
#include <utility>
#include <vector>

namespace rmx {

namespace {

// arcs whose ends share a module (of an out- or an in-CSR)
template <int G>
__global__ void k_intra_flags(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx, int32_t n,
                              const uint32_t *__restrict__ par, uint32_t *__restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n) return;
  if (optr[a + 1] - optr[a] > kBigRow) return;   // big rows: k_intra_flags_big
  const uint32_t pa = par[a];
  for (uint32_t k = optr[a] + lane; k < optr[a + 1]; k += G) flag[k] = par[oidx[k]] == pa ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_intra_flags_big(BigRows B, const uint32_t *__restrict__ optr,
                                                         const uint32_t *__restrict__ oidx,
                                                         const uint32_t *__restrict__ par, uint32_t *__restrict__ flag) {
  const uint32_t nch = *B.ctr;
  for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
    const uint32_t a = B.crow[c], b = B.cbeg[c];
    const uint32_t e = b + kBigRow < optr[a + 1] ? b + kBigRow : optr[a + 1];
    const uint32_t pa = par[a];
    for (uint32_t k = b + threadIdx.x; k < e; k += 256) flag[k] = par[oidx[k]] == pa ? 1u : 0u;
  }
}

__global__ void k_intra_ptr(const uint32_t *__restrict__ optr, int32_t n, int64_t m, const uint32_t *__restrict__ apos,
                            uint32_t total, uint32_t *__restrict__ sptr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  sptr[i] = (i < n && optr[i] < (uint64_t)m) ? apos[optr[i]] : total;
}

// the flagged arcs, compacted; a null weight destination is left out
__global__ void k_intra_copy(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ apos, int64_t m,
                             const uint32_t *__restrict__ oidx, const long long *__restrict__ wx,
                             uint32_t *__restrict__ sidx, long long *__restrict__ swx) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; k < m; k += stride) {
    if (!flag[k]) continue;
    const uint32_t j = apos[k];
    sidx[j] = oidx[k];
    swx[j] = wx[k];
  }
}

// unit statistics of the sub-networks: flow, teleport and dangling flow as
// in the network, out-flow over the intra-module arcs only
template <int G>
__global__ void k_sub_units(const uint32_t *__restrict__ sptr, int32_t n, ArcW aw, const Stat4 *__restrict__ us0,
                            Stat4 *__restrict__ us) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n) return;
  long long x = 0;
  const uint32_t e = sptr[a + 1];
  for (uint32_t k0 = sptr[a] + lane; k0 < e; k0 += kRowAhead * G) {
    long long v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? aw.fxw(k0 + j * G) : 0ll;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) x += v[j];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, G);
  if (lane == 0) {
    Stat4 s = us0[a];
    s.x = x;
    us[a] = s;
  }
}

__global__ void k_singleton_state(int32_t n, Stat4 *__restrict__ ms, const Stat4 *__restrict__ us,
                                  uint32_t *__restrict__ mod, int32_t *__restrict__ nmem, int32_t *__restrict__ nodes,
                                  int32_t *__restrict__ units, int32_t *__restrict__ stale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ms[i] = us[i];
  mod[i] = (uint32_t)i;
  nmem[i] = 1;
  nodes[i] = 1;
  units[i] = 1;
  stale[i] = 1;
}

// An instance whose sub-module partition ended above the codelength of its
// singletons returns to singletons.  The reference partitions an instance
// with sequential moves that each lower its codelength (move(),
// OmpRelaxmap.cpp:450), so it never ends above them; relaxed sub-sweeps can.
// Every module id of an instance is one of its own nodes (sub-module
// partitions offer no empty modules), so resetting the rows of its nodes
// resets all of its modules.
__global__ void k_keep_singletons_if_worse(int32_t n, const uint32_t *__restrict__ par,
                                           const double *__restrict__ L_single, const double *__restrict__ L,
                                           Stat4 *__restrict__ ms, const Stat4 *__restrict__ us,
                                           uint32_t *__restrict__ mod, int32_t *__restrict__ nmem,
                                           int32_t *__restrict__ nodes, int32_t *__restrict__ units,
                                           int32_t *__restrict__ stale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t I = par[i];
  if (!(L[I] > L_single[I])) return;
  ms[i] = us[i];
  mod[i] = (uint32_t)i;
  nmem[i] = 1;
  nodes[i] = 1;
  units[i] = 1;
  stale[i] = 1;
}

__global__ void k_heads(const uint32_t *__restrict__ key, int32_t n, uint32_t *__restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

__global__ void k_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, int32_t n,
                         uint32_t *__restrict__ start) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) start[pos[i]] = (uint32_t)i;
}

// Instances of more than kBigInst nodes are reduced by several workgroups,
// kBigInst nodes each, and their partial sums combined in chunk order; the
// others by one wavefront each.  The order of every sum is fixed by the
// instance's node list, so the results are reproducible.  The plan (which
// instances are big and where their chunks' partial sums go) is made once
// per coarse tune, since the instances do not change during its sweeps.
constexpr uint32_t kBigInst = 2048;
constexpr int kBigParts = 3;   // partial sums kept per chunk

__device__ inline void inst_range(const uint32_t *__restrict__ start, int32_t n_seg, int32_t n, int64_t s,
                                  uint32_t &b, uint32_t &e) {
  b = start[s];
  e = s + 1 < n_seg ? start[s + 1] : (uint32_t)n;
}

// one thread per instance: big ones get a list entry, consecutive chunk slots
// and the owner of every slot (ctr[0]: big instances, ctr[1]: chunk slots);
// the other instances of more than one node are listed in `multi` (ctr[2]):
// after the first totals only these and the big ones can change, since a
// one-node instance has no move to make
__global__ void k_big_plan(const uint32_t *__restrict__ start, int32_t n_seg, int32_t n, uint32_t *__restrict__ ctr,
                           uint32_t *__restrict__ list, uint32_t *__restrict__ base, uint32_t *__restrict__ owner,
                           uint32_t *__restrict__ multi) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  uint32_t b, e;
  inst_range(start, n_seg, n, s, b, e);
  if (e - b <= kBigInst) {
    if (e - b > 1) multi[atomicAdd(&ctr[2], 1u)] = (uint32_t)s;
    return;
  }
  const uint32_t nch = (e - b + kBigInst - 1) / kBigInst;
  const uint32_t li = atomicAdd(&ctr[0], 1u);
  const uint32_t sl = atomicAdd(&ctr[1], nch);
  list[li] = (uint32_t)s;
  base[li] = sl;
  for (uint32_t k = 0; k < nch; ++k) owner[sl + k] = li;
}

// the node range of chunk slot `slot`
__device__ inline void big_chunk(const BigInst &B, const uint32_t *__restrict__ start, int32_t n_seg, int32_t n,
                                 uint32_t slot, uint32_t &b, uint32_t &e) {
  const uint32_t li = B.owner[slot];
  uint32_t sb, se;
  inst_range(start, n_seg, n, B.list[li], sb, se);
  b = sb + (slot - B.base[li]) * kBigInst;
  e = b + kBigInst < se ? b + kBigInst : se;
}

// wavefront per instance: sum of phi(p) over its nodes (constant)
__global__ void k_inst_const(const uint32_t *__restrict__ start, int32_t n_seg, int32_t n,
                             const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ par,
                             const Stat4 *__restrict__ us, double *__restrict__ C) {
  const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (s >= n_seg) return;
  uint32_t b, e;
  inst_range(start, n_seg, n, s, b, e);
  if (e - b > kBigInst) return;   // k_inst_const_chunks / k_inst_const_big
  double a = 0.0;
  for (uint32_t i = b + lane; i < e; i += kWave) a += plogp(fx_to(us[nodes[i]].p));
  a = wave_sum(a);
  if (lane == 0) C[par[nodes[b]]] = a;
}

__global__ __launch_bounds__(256) void k_inst_const_chunks(BigInst B, uint32_t n_slots,
                                                           const uint32_t *__restrict__ start, int32_t n_seg,
                                                           int32_t n, const uint32_t *__restrict__ nodes,
                                                           const Stat4 *__restrict__ us) {
  __shared__ double lds[256 / kWave];
  for (uint32_t slot = blockIdx.x; slot < n_slots; slot += gridDim.x) {
    uint32_t b, e;
    big_chunk(B, start, n_seg, n, slot, b, e);
    double a = 0.0;
    for (uint32_t i = b + threadIdx.x; i < e; i += 256) a += plogp(fx_to(us[nodes[i]].p));
    a = block_sum<double, 256>(a, lds);
    if (threadIdx.x == 0) B.part[(size_t)slot * kBigParts] = a;
  }
}

__global__ void k_inst_const_big(BigInst B, uint32_t n_big, const uint32_t *__restrict__ start, int32_t n_seg,
                                 int32_t n, const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ par,
                                 double *__restrict__ C) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= n_big) return;
  uint32_t b, e;
  inst_range(start, n_seg, n, B.list[li], b, e);
  const uint32_t nch = (e - b + kBigInst - 1) / kBigInst;
  double a = 0.0;
  for (uint32_t k = 0; k < nch; ++k) a += B.part[(size_t)(B.base[li] + k) * kBigParts];
  C[par[nodes[b]]] = a;
}

// wavefront per instance: S = sum of its sub-module exits and its codelength
// L = phi(S) - 2 sum phi(q) + sum phi(q + P) - sum phi(p)  (Module.cpp:474-501)
__device__ inline void inst_terms(const Stat4 *__restrict__ ms, const int32_t *__restrict__ mnodes, uint32_t m,
                                  double &sq, double &pq, double &pqp) {
  if (mnodes[m] <= 0) return;
  const Stat4 st = ms[m];
  const double q = module_exit(st.p, st.t, st.d, st.x);
  sq += q;
  pq += plogp(q);
  pqp += plogp(q + fx_to(st.p));
}

// (one wavefront per listed instance, or per instance without a list)
__global__ void k_inst_totals(const uint32_t *__restrict__ start, int32_t n_seg, int32_t n,
                              const uint32_t *__restrict__ list, int64_t n_list,
                              const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ par,
                              const Stat4 *__restrict__ ms, const int32_t *__restrict__ mnodes,
                              const double *__restrict__ C, double *__restrict__ S_out, double *__restrict__ L_out) {
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wv >= n_list) return;
  const int64_t s = list ? (int64_t)list[wv] : wv;
  uint32_t b, e;
  inst_range(start, n_seg, n, s, b, e);
  if (e - b > kBigInst) return;   // k_inst_totals_chunks / k_inst_totals_big
  double sq = 0.0, pq = 0.0, pqp = 0.0;
  for (uint32_t i = b + lane; i < e; i += kWave) inst_terms(ms, mnodes, nodes[i], sq, pq, pqp);   // sub-module ids are node ids of the instance
  sq = wave_sum(sq);
  pq = wave_sum(pq);
  pqp = wave_sum(pqp);
  if (lane == 0) {
    const uint32_t I = par[nodes[b]];
    S_out[I] = sq;
    L_out[I] = plogp(sq) - 2.0 * pq + pqp - C[I];
  }
}

__global__ __launch_bounds__(256) void k_inst_totals_chunks(BigInst B, uint32_t n_slots,
                                                            const uint32_t *__restrict__ start, int32_t n_seg,
                                                            int32_t n, const uint32_t *__restrict__ nodes,
                                                            const Stat4 *__restrict__ ms,
                                                            const int32_t *__restrict__ mnodes) {
  __shared__ double lds[256 / kWave];
  for (uint32_t slot = blockIdx.x; slot < n_slots; slot += gridDim.x) {
    uint32_t b, e;
    big_chunk(B, start, n_seg, n, slot, b, e);
    double sq = 0.0, pq = 0.0, pqp = 0.0;
    for (uint32_t i = b + threadIdx.x; i < e; i += 256) inst_terms(ms, mnodes, nodes[i], sq, pq, pqp);
    sq = block_sum<double, 256>(sq, lds);
    pq = block_sum<double, 256>(pq, lds);
    pqp = block_sum<double, 256>(pqp, lds);
    if (threadIdx.x == 0) {
      double *o = B.part + (size_t)slot * kBigParts;
      o[0] = sq;
      o[1] = pq;
      o[2] = pqp;
    }
  }
}

__global__ void k_inst_totals_big(BigInst B, uint32_t n_big, const uint32_t *__restrict__ start, int32_t n_seg,
                                  int32_t n, const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ par,
                                  const double *__restrict__ C, double *__restrict__ S_out,
                                  double *__restrict__ L_out) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= n_big) return;
  uint32_t b, e;
  inst_range(start, n_seg, n, B.list[li], b, e);
  const uint32_t nch = (e - b + kBigInst - 1) / kBigInst;
  double sq = 0.0, pq = 0.0, pqp = 0.0;
  for (uint32_t k = 0; k < nch; ++k) {
    const double *o = B.part + (size_t)(B.base[li] + k) * kBigParts;
    sq += o[0];
    pq += o[1];
    pqp += o[2];
  }
  const uint32_t I = par[nodes[b]];
  S_out[I] = sq;
  L_out[I] = plogp(sq) - 2.0 * pq + pqp - C[I];
}

// instances with more than one member start active (in a batch of
// networks, only those of networks still in the super-step loop)
__global__ void k_inst_init(const uint32_t *__restrict__ start, int32_t n_seg, int32_t n,
                            const uint32_t *__restrict__ nodes, const uint32_t *__restrict__ par,
                            const uint8_t *__restrict__ bact, const uint32_t *__restrict__ bnode,
                            uint8_t *__restrict__ act) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const uint32_t b = start[s], e = s + 1 < n_seg ? start[s + 1] : (uint32_t)n;
  const uint32_t I = par[nodes[b]];
  act[I] = (e - b > 1 && (!bact || bact[bnode[I]])) ? 1 : 0;
}

// the stop rule of partition_module_network, per instance
// (OmpRelaxmap.cpp:458-459): stop when a sweep gains less than the threshold
__global__ void k_inst_update(const uint32_t *__restrict__ start, int32_t n_seg, const uint32_t *__restrict__ nodes,
                              const uint32_t *__restrict__ par, const double *__restrict__ Lprev,
                              const double *__restrict__ L, double thr_nats, uint8_t *__restrict__ act,
                              uint32_t *__restrict__ n_active) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg) return;
  const uint32_t I = par[nodes[start[s]]];
  if (!act[I]) return;
  if (Lprev[I] - L[I] < thr_nats) act[I] = 0;
  else atomicAdd(n_active, 1u);
}

__global__ void k_node_act(int32_t n, const uint32_t *__restrict__ par, const uint8_t *__restrict__ inst_act,
                           uint8_t *__restrict__ act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) act[i] = inst_act[par[i]];
}

__global__ void k_count_nonempty(const int32_t *__restrict__ nodes, int32_t n, uint32_t *__restrict__ cnt) {
  __shared__ uint32_t lds;
  if (threadIdx.x == 0) lds = 0;
  __syncthreads();
  uint32_t k = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    k += nodes[i] > 0 ? 1u : 0u;
  k = wave_sum(k);
  if ((threadIdx.x & (kWave - 1)) == 0 && k) atomicAdd(&lds, k);
  __syncthreads();
  if (threadIdx.x == 0 && lds) atomicAdd(cnt, lds);
}

// super nodes from sub-modules: exit flow from the super arcs, parent module,
// and the number of super nodes per module
template <int G>
__global__ void k_super_fix(const uint32_t *__restrict__ optr, const long long *__restrict__ owx, int32_t ns,
                            const uint32_t *__restrict__ par, Stat4 *__restrict__ us, uint32_t *__restrict__ smod,
                            int32_t *__restrict__ units) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= ns) return;
  long long x = 0;
  const uint32_t e = optr[a + 1];
  for (uint32_t k0 = optr[a] + lane; k0 < e; k0 += kRowAhead * G) {
    long long v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? owx[k0 + j * G] : 0ll;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) x += v[j];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, G);
  if (lane == 0) {
    us[a].x = x;
    const uint32_t parent = par[smod[a]];
    smod[a] = parent;
    atomicAdd(&units[parent], 1);
  }
}

// the sub-module state in the place of the network's node level and modules
void swap_in_sub(Ctx &c) {
  c.L0.swap(c.Lsub);
  c.ms.swap(c.sub_ms);
  c.mNodes.swap(c.sub_nodes);
  c.mUnits.swap(c.sub_units);
  c.mStale.swap(c.sub_stale);
}

}  // namespace

namespace {

BigInst big_view(Ctx &c) { return BigInst{c.ibig_list.p, c.ibig_base.p, c.ibig_owner.p, c.ibig_part.p}; }

unsigned big_grid(uint32_t slots) { return (unsigned)std::min<uint32_t>(slots, 2048u); }

// the chunk plan of the current instances (c.inst_start, c.n_inst)
void plan_big_instances(Ctx &c) {
  hipStream_t s = c.stream;
  const int64_t cap = (int64_t)c.N / kBigInst + 1;   // big instances; their chunks fit in 2 cap slots
  c.ibig_ctr.alloc(3);
  c.ibig_list.alloc(cap);
  c.ibig_base.alloc(cap);
  c.ibig_owner.alloc(2 * cap);
  c.ibig_part.alloc(2 * cap * kBigParts);
  c.inst_multi.alloc(std::max<int64_t>(c.n_inst, 1));
  fill_u32(c.ibig_ctr.p, 0u, 3, s);
  c.n_ibig = c.n_ibig_slots = c.n_multi = 0;
  if (c.n_inst <= 0) return;
  k_big_plan<<<grid_for(c.n_inst, 256), 256, 0, s>>>(c.inst_start.p, c.n_inst, c.N, c.ibig_ctr.p, c.ibig_list.p,
                                                     c.ibig_base.p, c.ibig_owner.p, c.inst_multi.p);
  RMX_LAUNCH_CHECK();
  uint32_t h[3] = {0, 0, 0};
  RMX_HIP(hipMemcpyAsync(h, c.ibig_ctr.p, sizeof(h), hipMemcpyDeviceToHost, s));
  c.sync();
  c.n_ibig = h[0];
  c.n_ibig_slots = h[1];
  c.n_multi = h[2];
}

}  // namespace

void instance_totals_async(Ctx &c, bool all) {
  if (c.n_inst <= 0) return;
  const int64_t nw = all ? c.n_inst : c.n_multi;   // wavefronts
  if (nw > 0) {
    k_inst_totals<<<grid_for(nw * kWave, 256), 256, 0, c.stream>>>(
        c.inst_start.p, c.n_inst, c.N, all ? nullptr : c.inst_multi.p, nw, c.inst_nodes.p, c.sub_par.p, c.ms.p,
        c.mNodes.p, c.inst_C.p, c.inst_S.p, c.inst_L.p);
    RMX_LAUNCH_CHECK();
  }
  if (c.n_ibig > 0) {
    k_inst_totals_chunks<<<big_grid(c.n_ibig_slots), 256, 0, c.stream>>>(big_view(c), c.n_ibig_slots, c.inst_start.p,
                                                                         c.n_inst, c.N, c.inst_nodes.p, c.ms.p,
                                                                         c.mNodes.p);
    RMX_LAUNCH_CHECK();
    k_inst_totals_big<<<grid_for(c.n_ibig, 256), 256, 0, c.stream>>>(big_view(c), c.n_ibig, c.inst_start.p, c.n_inst,
                                                                     c.N, c.inst_nodes.p, c.sub_par.p, c.inst_C.p,
                                                                     c.inst_S.p, c.inst_L.p);
    RMX_LAUNCH_CHECK();
  }
}

namespace {

// flags of the arcs of rows of more than kBigRow arcs (left out by k_intra_flags)
void intra_flags_big(Ctx &c, const uint32_t *ptr, const uint32_t *idx, int32_t n, int64_t m, uint32_t *flag) {
  DevBuf<uint32_t> bctr(c.cache), bcrow(c.cache), bcbeg(c.cache);
  DevBuf<long long> barcs(c.cache);
  bctr.alloc(1);
  barcs.alloc(1);
  bcrow.alloc(big_rows_cap(m));
  bcbeg.alloc(big_rows_cap(m));
  const BigRows big{bctr.p, bcrow.p, bcbeg.p, (unsigned long long *)barcs.p};
  plan_big_rows(ptr, n, big, c.stream);
  k_intra_flags_big<<<big_rows_grid(m), 256, 0, c.stream>>>(big, ptr, idx, c.sub_par.p, flag);
  RMX_LAUNCH_CHECK();
}   // the buffers go back to the cache: later work on the stream runs after the kernel

}  // namespace

int32_t generate_sub_modules(Ctx &c, double threshold, int max_iter, uint64_t seed) {
  hipStream_t s = c.stream;
  to_node_level(c);   // node modules current, level 0
  ensure_sweep_buffers(c);
  const int32_t n = c.N;
  Level &L0 = c.L0;
  const int64_t m = L0.m;

  // instance of every node: its module
  c.sub_par.alloc(n);
  RMX_HIP(hipMemcpyAsync(c.sub_par.p, L0.mod.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));

  // the intra-module level (generate_network_from_module): arcs kept in order
  Level &S = c.Lsub;
  S.n = n;
  uint32_t m_in = 0;
  {
    DevBuf<uint32_t> flag(c.cache), apos(c.cache), tot(c.cache);
    flag.alloc(m);
    apos.alloc(m);
    tot.alloc(1);
    if (m > 0) {
      k_intra_flags<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L0.out_ptr.p, L0.out_idx.p, n, c.sub_par.p,
                                                                        flag.p);
      RMX_LAUNCH_CHECK();
      intra_flags_big(c, L0.out_ptr.p, L0.out_idx.p, n, m, flag.p);
      exclusive_scan_u32(flag.p, apos.p, m, tot.p, c.scratch, s);
      RMX_HIP(hipMemcpyAsync(&m_in, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      c.sync();
    }
    S.m = m_in;
    S.out_ptr.alloc((int64_t)n + 1);
    S.out_idx.alloc(m_in);
    S.out_wx.alloc(m_in);   // the sweeps read only the fixed-point weights
    S.in_wx.alloc(m_in);
    k_intra_ptr<<<grid_for((int64_t)n + 1, 256), 256, 0, s>>>(L0.out_ptr.p, n, m, apos.p, m_in, S.out_ptr.p);
    RMX_LAUNCH_CHECK();
    if (m_in > 0) {
      k_intra_copy<<<(unsigned)std::min<int64_t>(16384, (m + 255) / 256), 256, 0, s>>>(
          flag.p, apos.p, m, L0.out_idx.p, L0.out_wx.p, S.out_idx.p, S.out_wx.p);
      RMX_LAUNCH_CHECK();
    }
    c.sync();
  }
  {   // in-lists: the node level's in-lists filtered the same way (their
      // ascending-tail order carries over, so no sort is needed)
    DevBuf<uint32_t> flag(c.cache), apos(c.cache), tot(c.cache);
    flag.alloc(m);
    apos.alloc(m);
    tot.alloc(1);
    uint32_t m_in2 = 0;
    if (m > 0) {
      k_intra_flags<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(L0.in_ptr.p, L0.in_idx.p, n, c.sub_par.p,
                                                                        flag.p);
      RMX_LAUNCH_CHECK();
      intra_flags_big(c, L0.in_ptr.p, L0.in_idx.p, n, m, flag.p);
      exclusive_scan_u32(flag.p, apos.p, m, tot.p, c.scratch, s);
      RMX_HIP(hipMemcpyAsync(&m_in2, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      c.sync();
    }
    if (m_in2 != m_in) throw HipError("intra-module in- and out-arcs disagree");
    S.in_ptr.alloc((int64_t)n + 1);
    S.in_idx.alloc(m_in);
    k_intra_ptr<<<grid_for((int64_t)n + 1, 256), 256, 0, s>>>(L0.in_ptr.p, n, m, apos.p, m_in, S.in_ptr.p);
    RMX_LAUNCH_CHECK();
    if (m_in > 0) {
      k_intra_copy<<<(unsigned)std::min<int64_t>(16384, (m + 255) / 256), 256, 0, s>>>(
          flag.p, apos.p, m, L0.in_idx.p, L0.in_wx.p, S.in_idx.p, S.in_wx.p);
      RMX_LAUNCH_CHECK();
    }
    c.sync();
  }
  build_buckets(c, S);
  S.us.alloc(n);
  S.nmem.alloc(n);
  S.mod.alloc(n);
  {   // out-flow of every unit over its intra arcs
    k_sub_units<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(S.out_ptr.p, n, arcw_out(c, S), L0.us.p, S.us.p);
    RMX_LAUNCH_CHECK();
  }
  // singleton sub-modules (Module(i, node), OmpRelaxmap.cpp:686-690)
  c.sub_ms.alloc(n);
  c.sub_nodes.alloc(n);
  c.sub_units.alloc(n);
  c.sub_stale.alloc(n);
  k_singleton_state<<<grid_for(n, 256), 256, 0, s>>>(n, c.sub_ms.p, S.us.p, S.mod.p, S.nmem.p, c.sub_nodes.p,
                                                     c.sub_units.p, c.sub_stale.p);
  RMX_LAUNCH_CHECK();

  // instances: nodes grouped by module (stable), one segment per module
  c.inst_nodes.alloc(n);
  c.inst_start.alloc(n);
  {
    DevBuf<uint32_t> key(c.cache), flag(c.cache), pos(c.cache), tot(c.cache);
    key.alloc(n);
    flag.alloc(n);
    pos.alloc(n);
    tot.alloc(1);
    RMX_HIP(hipMemcpyAsync(key.p, c.sub_par.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    iota_u32(c.inst_nodes.p, n, s);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
    radix_sort_pairs_u32(key.p, c.inst_nodes.p, n, 0, bits, c.scratch, s);
    k_heads<<<grid_for(n, 256), 256, 0, s>>>(key.p, n, flag.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(flag.p, pos.p, n, tot.p, c.scratch, s);
    k_starts<<<grid_for(n, 256), 256, 0, s>>>(flag.p, pos.p, n, c.inst_start.p);
    RMX_LAUNCH_CHECK();
    uint32_t ni = 0;
    RMX_HIP(hipMemcpyAsync(&ni, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    c.sync();
    c.n_inst = (int32_t)ni;
  }
  c.inst_S.alloc(n);
  c.inst_L.alloc(n);
  c.inst_Lprev.alloc(n);
  c.inst_C.alloc(n);
  c.inst_act.alloc(n);
  const unsigned gw = grid_for((int64_t)c.n_inst * kWave, 256), gs = grid_for(c.n_inst, 256);
  plan_big_instances(c);
  k_inst_const<<<gw, 256, 0, s>>>(c.inst_start.p, c.n_inst, n, c.inst_nodes.p, c.sub_par.p, S.us.p, c.inst_C.p);
  RMX_LAUNCH_CHECK();
  if (c.n_ibig > 0) {
    k_inst_const_chunks<<<big_grid(c.n_ibig_slots), 256, 0, s>>>(big_view(c), c.n_ibig_slots, c.inst_start.p, c.n_inst,
                                                                 n, c.inst_nodes.p, S.us.p);
    RMX_LAUNCH_CHECK();
    k_inst_const_big<<<grid_for(c.n_ibig, 256), 256, 0, s>>>(big_view(c), c.n_ibig, c.inst_start.p, c.n_inst, n,
                                                             c.inst_nodes.p, c.sub_par.p, c.inst_C.p);
    RMX_LAUNCH_CHECK();
  }
  fill_u8(c.inst_act.p, 0, n, s);
  k_inst_init<<<gs, 256, 0, s>>>(c.inst_start.p, c.n_inst, n, c.inst_nodes.p, c.sub_par.p,
                                 c.nb > 0 ? c.b_step.p : nullptr, c.b_node.p, c.inst_act.p);
  RMX_LAUNCH_CHECK();

  // partition_module_network on every instance (OmpRelaxmap.cpp:436-498,
  // fast = true: no super-node phase)
  swap_in_sub(c);
  c.sub_mode = 1;
  const double thr_nats = threshold * std::log(2.0);
  DevBuf<double> L_single(c.cache);   // each instance's codelength as singletons
  L_single.alloc(n);
  try {
    instance_totals_async(c, true);
    RMX_HIP(hipMemcpyAsync(L_single.p, c.inst_L.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_node_act<<<grid_for(n, 256), 256, 0, s>>>(n, c.sub_par.p, c.inst_act.p, c.act.p);
    RMX_LAUNCH_CHECK();
    fill_u8(c.act_next.p, 0, n, s);
    for (int it = 0; it < max_iter; ++it) {
      RMX_HIP(hipMemcpyAsync(c.inst_Lprev.p, c.inst_L.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      int64_t vis = 0;
      move_sweep(c, true, 0.0, seed ^ (0x2545f4914f6cdd1dull * (uint64_t)(it + 1)), &vis);
      c.timing.edge_visits -= vis;   // not part of the reported sweep throughput
      c.timing.sweeps--;
      // the sweep left its frontier in act: the instances' own state replaces it
      DevBuf<uint32_t> cnt(c.cache);
      cnt.alloc(1);
      fill_u32(cnt.p, 0u, 1, s);
      k_inst_update<<<gs, 256, 0, s>>>(c.inst_start.p, c.n_inst, c.inst_nodes.p, c.sub_par.p, c.inst_Lprev.p,
                                       c.inst_L.p, thr_nats, c.inst_act.p, cnt.p);
      RMX_LAUNCH_CHECK();
      k_node_act<<<grid_for(n, 256), 256, 0, s>>>(n, c.sub_par.p, c.inst_act.p, c.act.p);
      RMX_LAUNCH_CHECK();
      uint32_t active = 0;
      RMX_HIP(hipMemcpyAsync(&active, cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      c.sync();
      if (active == 0) break;
    }
  } catch (...) {
    c.sub_mode = 0;
    swap_in_sub(c);
    throw;
  }
  c.sub_mode = 0;
  swap_in_sub(c);
  reset_active(c);
  k_keep_singletons_if_worse<<<grid_for(n, 256), 256, 0, s>>>(n, c.sub_par.p, L_single.p, c.inst_L.p, c.sub_ms.p,
                                                              S.us.p, S.mod.p, S.nmem.p, c.sub_nodes.p,
                                                              c.sub_units.p, c.sub_stale.p);
  RMX_LAUNCH_CHECK();

  // number of sub-modules
  DevBuf<uint32_t> cnt(c.cache);
  cnt.alloc(1);
  fill_u32(cnt.p, 0u, 1, s);
  k_count_nonempty<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, s>>>(c.sub_nodes.p, n, cnt.p);
  RMX_LAUNCH_CHECK();
  uint32_t ns = 0;
  RMX_HIP(hipMemcpyAsync(&ns, cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  c.sync();
  c.n_sub = (int32_t)ns;
  return c.n_sub;
}

void contract_sub_modules(Ctx &c) {
  if (c.n_sub <= 0) throw HipError("generate sub-modules first");
  hipStream_t s = c.stream;
  to_node_level(c);
  // contract by the sub-module labels and statistics: the node level's
  // module labels and module rows are exchanged for the sub-modules' for the
  // duration of the contraction
  c.L0.mod.swap(c.Lsub.mod);
  c.ms.swap(c.sub_ms);
  c.mNodes.swap(c.sub_nodes);
  try {
    contract(c);
  } catch (...) {
    c.L0.mod.swap(c.Lsub.mod);
    c.ms.swap(c.sub_ms);
    c.mNodes.swap(c.sub_nodes);
    throw;
  }
  c.L0.mod.swap(c.Lsub.mod);
  c.ms.swap(c.sub_ms);
  c.mNodes.swap(c.sub_nodes);
  // super nodes: exit flow over the super arcs, parent module, units per module
  Level &L1 = c.L1;
  fill_u32((uint32_t *)c.mUnits.p, 0u, c.N, s);
  if (L1.n > 0) {
    k_super_fix<16><<<grid_for((int64_t)L1.n * 16, 256), 256, 0, s>>>(L1.out_ptr.p, L1.out_wx.p, L1.n, c.sub_par.p,
                                                                      L1.us.p, L1.mod.p, c.mUnits.p);
    RMX_LAUNCH_CHECK();
  }
  // node modules are the parents again (contract() relabelled node_sp only)
  c.sync();
  c.n_sub = 0;   // consumed
}

}  // namespace rmx
