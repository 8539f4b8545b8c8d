// Map-equation codelength of the current partition.
//
// Tracked form (the sweeps): module statistics P, T, D (flow, teleport,
// dangling flow) and X (flow on arcs leaving the module) are fixed-point
// sums; q_M = (alpha P + beta D)(1 - T) + beta X is the exit probability
// of Module.cpp:2978-2991, and
//   L = phi(sum q) - 2 sum phi(q) + sum phi(q + P) - sum_a phi(p_a)
// as in Network::calibrate (Module.cpp:474-501).
//
// From-scratch form (Network::calculateCodeLength, Module.cpp:2966-3002):
// every node's exit-arc flow, then nodes stably sorted by module and one
// wavefront per module summing its members in FP64 (a module of more than
// 2048 members: one workgroup per chunk of them, the chunks added in order).
// The tree order is fixed, so the result is reproducible; it differs from the
// reference's sequential order by rounding only.
#include "devutil.h"

namespace rmx {

namespace {

constexpr int kRedGrid = 1024;
constexpr int kRedBlock = 256;
enum { kS = 0, kR = 1, kD = 2, kL = 3, kNM = 7 };

__device__ inline double reduce_partials(const double *__restrict__ part, double *lds) {
  double a = 0.0;
  for (int i = threadIdx.x; i < kRedGrid; i += kRedBlock) a += part[i];
  a = block_sum<double, kRedBlock>(a, lds);
  if (threadIdx.x == 0) lds[0] = a;
  __syncthreads();
  double s = lds[0];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(kRedBlock) void k_module_totals(const Stat4 *__restrict__ ms,
                                                             const int32_t *__restrict__ mNodes, int32_t n,
                                                             double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  double s = 0, se = 0, ss = 0, cnt = 0;
  for (int64_t m = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; m < n; m += (int64_t)kRedGrid * kRedBlock) {
    if (mNodes[m] > 0) {
      const Stat4 st = ms[m];
      double q = module_exit(st.p, st.t, st.d, st.x);
      s += q;
      se += plogp(q);
      ss += plogp(q + fx_to(st.p));
      cnt += 1.0;
    }
  }
  s = block_sum<double, kRedBlock>(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  se = block_sum<double, kRedBlock>(se, lds);
  if (threadIdx.x == 0) part[kRedGrid + blockIdx.x] = se;
  ss = block_sum<double, kRedBlock>(ss, lds);
  if (threadIdx.x == 0) part[2 * kRedGrid + blockIdx.x] = ss;
  cnt = block_sum<double, kRedBlock>(cnt, lds);
  if (threadIdx.x == 0) part[3 * kRedGrid + blockIdx.x] = cnt;
}

// qv[m] = q of module m, 0 for an empty module (its statistics are then
// exactly zero)
__global__ void k_qv_all(const Stat4 *__restrict__ ms, const int32_t *__restrict__ mNodes, int32_t n,
                         double *__restrict__ qv) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double q = 0.0;
  if (mNodes[m] > 0) {
    const Stat4 st = ms[m];
    q = module_exit(st.p, st.t, st.d, st.x);
  }
  qv[m] = q;
}

// S in k_module_totals' launch shape and order: an empty module adds +0.0,
// which leaves every partial sum unchanged
__global__ __launch_bounds__(kRedBlock) void k_sum_qv(const double *__restrict__ qv, int32_t n,
                                                      double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  constexpr int64_t kStride = (int64_t)kRedGrid * kRedBlock;
  double s = 0;
  // four loads in flight per lane; the additions keep the one-at-a-time order
  int64_t m = (int64_t)blockIdx.x * kRedBlock + threadIdx.x;
  for (; m + 3 * kStride < n; m += 4 * kStride) {
    const double v0 = qv[m], v1 = qv[m + kStride], v2 = qv[m + 2 * kStride], v3 = qv[m + 3 * kStride];
    s += v0;
    s += v1;
    s += v2;
    s += v3;
  }
  for (; m < n; m += kStride) s += qv[m];
  s = block_sum<double, kRedBlock>(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void k_sum_final(const double *__restrict__ part, double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  const double s = reduce_partials(part, lds);
  if (threadIdx.x == 0) scal[kS] = s;
}

// sum of phi(p) in the launch shape and order of the initial calibrate's
// (k_calib in flow.hip: a grid-stride sum per thread, then the block and
// partial trees), so the value equals the one the context keeps
__global__ __launch_bounds__(kRedBlock) void k_plogp_partials(const double *__restrict__ p, int32_t n,
                                                              double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  double sp = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock)
    sp += plogp(p[i]);
  sp = block_sum<double, kRedBlock>(sp, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = sp;
}

// k_totals_final with sum phi(p) from device partials
__global__ __launch_bounds__(kRedBlock) void k_totals_final_dev(const double *__restrict__ part,
                                                                const double *__restrict__ sp_part,
                                                                double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  double s = reduce_partials(part, lds);
  double se = reduce_partials(part + kRedGrid, lds);
  double ss = reduce_partials(part + 2 * kRedGrid, lds);
  double cnt = reduce_partials(part + 3 * kRedGrid, lds);
  double sp = reduce_partials(sp_part, lds);
  if (threadIdx.x == 0) {
    scal[kS] = s;
    scal[kL] = plogp(s) - 2.0 * se + ss - sp;
    scal[kNM] = cnt;
  }
}

__global__ __launch_bounds__(kRedBlock) void k_totals_final(const double *__restrict__ part, double all_log_all,
                                                            double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  double s = reduce_partials(part, lds);
  double se = reduce_partials(part + kRedGrid, lds);
  double ss = reduce_partials(part + 2 * kRedGrid, lds);
  double cnt = reduce_partials(part + 3 * kRedGrid, lds);
  if (threadIdx.x == 0) {
    scal[kS] = s;
    scal[kL] = plogp(s) - 2.0 * se + ss - all_log_all;
    scal[kNM] = cnt;
  }
}

// A batch: one workgroup per instance over its module ids (modules never
// leave their instance).  S = sum of exits, L = phi(S) - 2 sum phi(q) +
// sum phi(q + P) - sum phi(p) (Module.cpp:474-501 per instance), and the
// number of non-empty modules.
// Per-instance totals in two fixed-order steps: one workgroup per chunk of
// at most kTotChunk module ids (a chunk never spans two instances, so a
// large instance is spread over many workgroups), then one thread per
// instance adding its chunks in order.
__global__ __launch_bounds__(kRedBlock) void k_chunk_totals(const uint32_t *__restrict__ cbeg,
                                                            const uint32_t *__restrict__ cend,
                                                            const Stat4 *__restrict__ ms,
                                                            const int32_t *__restrict__ mNodes,
                                                            double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  const uint32_t b = cbeg[blockIdx.x], e = cend[blockIdx.x];
  double sq = 0, pq = 0, pqp = 0, cnt = 0;
  for (uint32_t m = b + threadIdx.x; m < e; m += kRedBlock) {
    if (mNodes[m] > 0) {
      const Stat4 st = ms[m];
      const double q = module_exit(st.p, st.t, st.d, st.x);
      sq += q;
      pq += plogp(q);
      pqp += plogp(q + fx_to(st.p));
      cnt += 1.0;
    }
  }
  sq = block_sum<double, kRedBlock>(sq, lds);
  pq = block_sum<double, kRedBlock>(pq, lds);
  pqp = block_sum<double, kRedBlock>(pqp, lds);
  cnt = block_sum<double, kRedBlock>(cnt, lds);
  if (threadIdx.x == 0) {
    double *o = part + 4 * (int64_t)blockIdx.x;
    o[0] = sq;
    o[1] = pq;
    o[2] = pqp;
    o[3] = cnt;
  }
}

__global__ void k_batch_totals(int32_t nb, const uint32_t *__restrict__ ichunk, const double *__restrict__ part,
                               const double *__restrict__ C, double *__restrict__ S_out, double *__restrict__ L_out,
                               int32_t *__restrict__ nmod) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  double sq = 0, pq = 0, pqp = 0, cnt = 0;
  for (uint32_t k = ichunk[I]; k < ichunk[I + 1]; ++k) {
    const double *o = part + 4 * (int64_t)k;
    sq += o[0];
    pq += o[1];
    pqp += o[2];
    cnt += o[3];
  }
  S_out[I] = sq;
  L_out[I] = plogp(sq) - 2.0 * pq + pqp - C[I];
  nmod[I] = (int32_t)cnt;
}

// sums over the instances (the host's view of a batch: total codelength,
// total exit, modules)
__global__ __launch_bounds__(kRedBlock) void k_batch_sum(int32_t nb, const double *__restrict__ S,
                                                         const double *__restrict__ L,
                                                         const int32_t *__restrict__ nmod, double *__restrict__ scal) {
  __shared__ double lds[kRedBlock / kWave];
  double s = 0, l = 0, k = 0;
  for (int32_t I = threadIdx.x; I < nb; I += kRedBlock) {
    s += S[I];
    l += L[I];
    k += (double)nmod[I];
  }
  s = block_sum<double, kRedBlock>(s, lds);
  l = block_sum<double, kRedBlock>(l, lds);
  k = block_sum<double, kRedBlock>(k, lds);
  if (threadIdx.x == 0) {
    scal[kS] = s;
    scal[kL] = l;
    scal[kNM] = k;
  }
}

__global__ void k_zero_modules(int32_t n, Stat4 *ms, int32_t *mNodes, int32_t *mUnits) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ms[i] = Stat4{0, 0, 0, 0};
  mNodes[i] = 0; mUnits[i] = 0;
}

__global__ void k_accumulate_units(int32_t n, const uint32_t *__restrict__ mod, const Stat4 *__restrict__ us,
                                   const int32_t *__restrict__ nmem, Stat4 *ms, int32_t *mNodes, int32_t *mUnits) {
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint32_t m = mod[u];
  const Stat4 st = us[u];
  atomicAdd((unsigned long long *)&ms[m].p, (unsigned long long)st.p);
  atomicAdd((unsigned long long *)&ms[m].t, (unsigned long long)st.t);
  atomicAdd((unsigned long long *)&ms[m].d, (unsigned long long)st.d);
  atomicAdd(&mNodes[m], nmem[u]);
  atomicAdd(&mUnits[m], 1);
}

// X_M = sum over arcs a->b with a in M, b not in M of the arc flow
template <int G>
__global__ void k_exit_flows(int32_t n, const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx,
                             ArcW aw, const uint32_t *__restrict__ mod, Stat4 *ms) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t u = t / G;
  const int lane = (int)(t % G);
  if (u >= n) return;
  const uint32_t m = mod[u];
  long long acc = 0;
  const uint32_t e = optr[u + 1];
  for (uint32_t k0 = optr[u] + lane; k0 < e; k0 += kRowAhead * G) {
    uint32_t v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? oidx[k0 + j * G] : 0u;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? mod[v[j]] : m;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j)
      if (v[j] != m) acc += aw.fxw(k0 + j * G);
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
  if (lane == 0 && acc != 0) atomicAdd((unsigned long long *)&ms[m].x, (unsigned long long)acc);
}

__global__ void k_zero_exit(int32_t n, Stat4 *ms) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ms[i].x = 0;
}

template <typename T>
__global__ void k_copy(const T *__restrict__ a, T *__restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_seg_heads(const uint32_t *__restrict__ key, int32_t n, uint32_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

__global__ void k_seg_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, int32_t n,
                             uint32_t *__restrict__ start) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) start[pos[i]] = (uint32_t)i;
}

// flow on each node's out-arcs that leave its module (FP64)
template <int G>
__global__ void k_p5_ext(int32_t n, const uint32_t *__restrict__ optr, const uint32_t *__restrict__ oidx,
                         const double *__restrict__ ow, const uint32_t *__restrict__ mod, double *__restrict__ ext) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t a = t / G;
  const int lane = (int)(t % G);
  if (a >= n) return;
  const uint32_t m = mod[a];
  double acc = 0.0;
  const uint32_t e = optr[a + 1];
  for (uint32_t k0 = optr[a] + lane; k0 < e; k0 += kRowAhead * G) {
    uint32_t v[kRowAhead];
    double w[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? oidx[k0 + j * G] : 0u;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) {
      w[j] = k0 + j * G < e ? ow[k0 + j * G] : 0.0;
      v[j] = k0 + j * G < e ? mod[v[j]] : m;
    }
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j)   // in the plain loop's order
      if (v[j] != m) acc += w[j];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
  if (lane == 0) ext[a] = acc;
}

// one wavefront per module (its members are a segment of the module-sorted
// node list): FP64 sums of flow, teleport, dangling flow and exit-arc flow,
// then the module exit of Module.cpp:2978-2991 in the reference's form.
// Modules of more than kP5Big members are summed in chunks of kP5Big members
// by k_p5_chunks (one workgroup per chunk) and combined in chunk order by
// k_p5_big, so that one large module does not hold a single wavefront for
// the whole launch; every sum keeps an order fixed by the member list.
constexpr uint32_t kP5Big = 2048;

__device__ inline void p5_seg(const uint32_t *__restrict__ start, uint32_t nseg, int32_t n, uint32_t s,
                              uint32_t &b, uint32_t &e) {
  b = start[s];
  e = (s + 1 < nseg) ? start[s + 1] : (uint32_t)n;
}

__device__ inline void p5_exit(double P, double T, double D, double X, double &q, double &qp) {
  q = kAlpha * (1.0 - T) * P;
  q += kBeta * (X + (1.0 - T) * D);
  qp = q + P;
}

__global__ void k_p5_modules(const uint32_t *__restrict__ start, const uint32_t *__restrict__ nseg_p, int32_t n,
                             const uint32_t *__restrict__ node, const double *__restrict__ p,
                             const double *__restrict__ tau, const uint8_t *__restrict__ dang,
                             const double *__restrict__ ext, double *__restrict__ segq, double *__restrict__ segqp) {
  const uint32_t nseg = *nseg_p;
  const int64_t sidx = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (sidx >= nseg) return;
  uint32_t b, e;
  p5_seg(start, nseg, n, (uint32_t)sidx, b, e);
  if (e - b > kP5Big) return;   // k_p5_chunks / k_p5_big
  double P = 0.0, T = 0.0, D = 0.0, X = 0.0;
  for (uint32_t i = b + lane; i < e; i += kWave) {
    const uint32_t a = node[i];
    P += p[a];
    T += tau[a];
    if (dang[a]) D += p[a];
    X += ext[a];
  }
  P = wave_sum(P);
  T = wave_sum(T);
  D = wave_sum(D);
  X = wave_sum(X);
  if (lane == 0) p5_exit(P, T, D, X, segq[sidx], segqp[sidx]);
}

// Chunk plan of the large modules: ctr[0] large modules, ctr[1] chunk slots;
// list / base per large module, owner per slot.
__global__ void k_p5_plan(const uint32_t *__restrict__ start, const uint32_t *__restrict__ nseg_p, int32_t n,
                          uint32_t *__restrict__ ctr, uint32_t *__restrict__ list, uint32_t *__restrict__ base,
                          uint32_t *__restrict__ owner) {
  const uint32_t nseg = *nseg_p;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  uint32_t b, e;
  p5_seg(start, nseg, n, (uint32_t)s, b, e);
  if (e - b <= kP5Big) return;
  const uint32_t nch = (e - b + kP5Big - 1) / kP5Big;
  const uint32_t li = atomicAdd(&ctr[0], 1u);
  const uint32_t sl = atomicAdd(&ctr[1], nch);
  list[li] = (uint32_t)s;
  base[li] = sl;
  for (uint32_t k = 0; k < nch; ++k) owner[sl + k] = li;
}

__global__ __launch_bounds__(256) void k_p5_chunks(const uint32_t *__restrict__ ctr, const uint32_t *__restrict__ list,
                                                   const uint32_t *__restrict__ base,
                                                   const uint32_t *__restrict__ owner, const uint32_t *__restrict__ start,
                                                   const uint32_t *__restrict__ nseg_p, int32_t n,
                                                   const uint32_t *__restrict__ node, const double *__restrict__ p,
                                                   const double *__restrict__ tau, const uint8_t *__restrict__ dang,
                                                   const double *__restrict__ ext, double *__restrict__ part) {
  __shared__ double lds[256 / kWave];
  const uint32_t nseg = *nseg_p, slots = ctr[1];
  for (uint32_t slot = blockIdx.x; slot < slots; slot += gridDim.x) {
    const uint32_t li = owner[slot];
    uint32_t sb, se;
    p5_seg(start, nseg, n, list[li], sb, se);
    const uint32_t b = sb + (slot - base[li]) * kP5Big, e = b + kP5Big < se ? b + kP5Big : se;
    double P = 0.0, T = 0.0, D = 0.0, X = 0.0;
    for (uint32_t i = b + threadIdx.x; i < e; i += 256) {
      const uint32_t a = node[i];
      P += p[a];
      T += tau[a];
      if (dang[a]) D += p[a];
      X += ext[a];
    }
    P = block_sum<double, 256>(P, lds);
    T = block_sum<double, 256>(T, lds);
    D = block_sum<double, 256>(D, lds);
    X = block_sum<double, 256>(X, lds);
    if (threadIdx.x == 0) {
      double *o = part + 4 * (size_t)slot;
      o[0] = P;
      o[1] = T;
      o[2] = D;
      o[3] = X;
    }
  }
}

__global__ void k_p5_big(const uint32_t *__restrict__ ctr, const uint32_t *__restrict__ list,
                         const uint32_t *__restrict__ base, const uint32_t *__restrict__ start,
                         const uint32_t *__restrict__ nseg_p, int32_t n, const double *__restrict__ part,
                         double *__restrict__ segq, double *__restrict__ segqp) {
  const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= ctr[0]) return;
  const uint32_t s = list[li];
  uint32_t b, e;
  p5_seg(start, *nseg_p, n, s, b, e);
  const uint32_t nch = (e - b + kP5Big - 1) / kP5Big;
  double P = 0.0, T = 0.0, D = 0.0, X = 0.0;
  for (uint32_t k = 0; k < nch; ++k) {
    const double *o = part + 4 * (size_t)(base[li] + k);
    P += o[0];
    T += o[1];
    D += o[2];
    X += o[3];
  }
  p5_exit(P, T, D, X, segq[s], segqp[s]);
}

__global__ __launch_bounds__(kRedBlock) void k_p5_partials(const double *__restrict__ segq,
                                                           const double *__restrict__ segqp,
                                                           const uint32_t *__restrict__ nseg_p,
                                                           double *__restrict__ part) {
  __shared__ double lds[kRedBlock / kWave];
  const uint32_t nseg = *nseg_p;
  double s = 0, se = 0, ss = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < nseg; i += (int64_t)kRedGrid * kRedBlock) {
    s += segq[i];
    se += plogp(segq[i]);
    ss += plogp(segqp[i]);
  }
  s = block_sum<double, kRedBlock>(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  se = block_sum<double, kRedBlock>(se, lds);
  if (threadIdx.x == 0) part[kRedGrid + blockIdx.x] = se;
  ss = block_sum<double, kRedBlock>(ss, lds);
  if (threadIdx.x == 0) part[2 * kRedGrid + blockIdx.x] = ss;
  if (threadIdx.x == 0) part[3 * kRedGrid + blockIdx.x] = 0.0;
}

}  // namespace

ArcW arcw_out(const Ctx &, const Level &L) { return ArcW{L.out_wx.p}; }

ArcW arcw_in(const Ctx &, const Level &L) { return ArcW{L.in_wx.p}; }

void module_totals_async(Ctx &c) {
  if (c.nb > 0) {   // per instance: S for the sweeps, L for the driver
    c.b_S.alloc(c.nb);
    c.b_L.alloc(c.nb);
    c.b_nmod.alloc(c.nb);
    if (c.n_chunks > 0) {
      k_chunk_totals<<<c.n_chunks, kRedBlock, 0, c.stream>>>(c.b_cbeg.p, c.b_cend.p, c.ms.p, c.mNodes.p,
                                                             c.b_cpart.p);
      RMX_LAUNCH_CHECK();
    }
    k_batch_totals<<<grid_for(c.nb, 256), 256, 0, c.stream>>>(c.nb, c.b_ichunk.p, c.b_cpart.p, c.b_C.p, c.b_S.p,
                                                               c.b_L.p, c.b_nmod.p);
    RMX_LAUNCH_CHECK();
    k_batch_sum<<<1, kRedBlock, 0, c.stream>>>(c.nb, c.b_S.p, c.b_L.p, c.b_nmod.p, c.scal.p);
    RMX_LAUNCH_CHECK();
    return;
  }
  k_module_totals<<<kRedGrid, kRedBlock, 0, c.stream>>>(c.ms.p, c.mNodes.p, c.N,
                                                         c.partial.p);
  RMX_LAUNCH_CHECK();
  k_totals_final<<<1, kRedBlock, 0, c.stream>>>(c.partial.p, c.all_log_all, c.scal.p);
  RMX_LAUNCH_CHECK();
}

void exit_probs_refresh(Ctx &c) {
  c.qv.alloc(c.N);
  k_qv_all<<<grid_for(c.N, 256), 256, 0, c.stream>>>(c.ms.p, c.mNodes.p, c.N, c.qv.p);
  RMX_LAUNCH_CHECK();
}

void sum_exit_async(Ctx &c) {
  k_sum_qv<<<kRedGrid, kRedBlock, 0, c.stream>>>(c.qv.p, c.N, c.partial.p);
  RMX_LAUNCH_CHECK();
  k_sum_final<<<1, kRedBlock, 0, c.stream>>>(c.partial.p, c.scal.p);
  RMX_LAUNCH_CHECK();
}

void module_totals(Ctx &c, double *S, double *L) {
  module_totals_async(c);
  RMX_HIP(hipMemcpyAsync(c.h_scal, c.scal.p, sizeof(double) * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  if (S) *S = c.h_scal[kS];
  if (L) *L = c.h_scal[kL];
  c.n_module = (int32_t)c.h_scal[kNM];
}

void rebuild_exit_flows(Ctx &c) {
  Level &L = c.cur();
  k_zero_exit<<<grid_for(c.N, 256), 256, 0, c.stream>>>(c.N, c.ms.p);
  RMX_LAUNCH_CHECK();
  k_exit_flows<16><<<grid_for((int64_t)L.n * 16, 256), 256, 0, c.stream>>>(L.n, L.out_ptr.p, L.out_idx.p,
                                                                            arcw_out(c, L), L.mod.p, c.ms.p);
  RMX_LAUNCH_CHECK();
}

void set_partition(Ctx &c, const uint32_t *d_part) {
  if (c.level != 0) throw HipError("set_partition needs the node level");
  Level &L = c.L0;
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  RMX_HIP(hipMemcpyAsync(L.mod.p, d_part, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  unsigned g = grid_for(n, 256);
  k_zero_modules<<<g, 256, 0, s>>>(n, c.ms.p, c.mNodes.p, c.mUnits.p);
  RMX_LAUNCH_CHECK();
  k_accumulate_units<<<g, 256, 0, s>>>(n, L.mod.p, L.us.p, L.nmem.p, c.ms.p, c.mNodes.p, c.mUnits.p);
  RMX_LAUNCH_CHECK();
  rebuild_exit_flows(c);
  k_copy<int32_t><<<g, 256, 0, s>>>(c.mNodes.p, c.mStale.p, n);
  RMX_LAUNCH_CHECK();
  module_totals(c, &c.sum_exit, &c.codelength);
}

void p5_launch(int32_t n, const uint32_t *optr, const uint32_t *oidx, const double *ow, const uint32_t *mod,
               const double *p, const double *tau, const uint8_t *dang, double all_log_all, bool device_log_all,
               const P5Work &w, Scratch &scratch, hipStream_t s) {
  RMX_HIP(hipMemcpyAsync(w.key, mod, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  iota_u32(w.node, n, s);
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
  radix_sort_pairs_u32(w.key, w.node, n, 0, bits, scratch, s);
  unsigned g = grid_for(n, 256);
  k_seg_heads<<<g, 256, 0, s>>>(w.key, n, w.flag);
  RMX_LAUNCH_CHECK();
  exclusive_scan_u32(w.flag, w.pos, n, w.total, scratch, s);
  k_seg_starts<<<g, 256, 0, s>>>(w.flag, w.pos, n, w.start);
  RMX_LAUNCH_CHECK();
  k_p5_ext<16><<<grid_for((int64_t)n * 16, 256), 256, 0, s>>>(n, optr, oidx, ow, mod, w.ext);
  RMX_LAUNCH_CHECK();
  k_p5_modules<<<grid_for((int64_t)n * kWave, 256), 256, 0, s>>>(w.start, w.total, n, w.node, p, tau, dang, w.ext,
                                                                 w.segq, w.segqp);
  RMX_LAUNCH_CHECK();
  if ((uint32_t)n > kP5Big) {
    // modules of more than kP5Big members: the chunk plan lives in the
    // node-sized temporaries that the segmentation no longer needs (flag:
    // counters, list, base, owner; pos: partial sums), at most n / kP5Big + 1
    // large modules and twice that many chunks
    const int64_t cap = (int64_t)n / kP5Big + 1;
    uint32_t *ctr = w.flag, *list = ctr + 2, *base = list + cap, *owner = base + cap;
    double *part = (double *)w.pos;
    fill_u32(ctr, 0u, 2, s);
    k_p5_plan<<<g, 256, 0, s>>>(w.start, w.total, n, ctr, list, base, owner);
    RMX_LAUNCH_CHECK();
    k_p5_chunks<<<(unsigned)std::min<int64_t>(2 * cap, 2048), 256, 0, s>>>(ctr, list, base, owner, w.start, w.total,
                                                                            n, w.node, p, tau, dang, w.ext, part);
    RMX_LAUNCH_CHECK();
    k_p5_big<<<grid_for(cap, 256), 256, 0, s>>>(ctr, list, base, w.start, w.total, n, part, w.segq, w.segqp);
    RMX_LAUNCH_CHECK();
  }
  k_p5_partials<<<kRedGrid, kRedBlock, 0, s>>>(w.segq, w.segqp, w.total, w.partial);
  RMX_LAUNCH_CHECK();
  if (device_log_all) {
    k_plogp_partials<<<kRedGrid, kRedBlock, 0, s>>>(p, n, w.partial + 4 * kRedGrid);
    RMX_LAUNCH_CHECK();
    k_totals_final_dev<<<1, kRedBlock, 0, s>>>(w.partial, w.partial + 4 * kRedGrid, w.scal);
  } else {
    k_totals_final<<<1, kRedBlock, 0, s>>>(w.partial, all_log_all, w.scal);
  }
  RMX_LAUNCH_CHECK();
}

double calculate_codelength(Ctx &c) {
  Level &L = c.L0;
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  DevBuf<uint32_t> key(c.cache), node(c.cache), flag(c.cache), pos(c.cache), start(c.cache), total(c.cache);
  DevBuf<double> segq(c.cache), segqp(c.cache), ext(c.cache);
  key.alloc(n);
  ext.alloc(n);
  node.alloc(n);
  flag.alloc(n);
  pos.alloc(n);
  start.alloc(n);
  total.alloc(1);
  segq.alloc(n);
  segqp.alloc(n);
  P5Work w{key.p, node.p, flag.p, pos.p, start.p, total.p, ext.p, segq.p, segqp.p, c.partial.p, c.scal.p + 8};
  p5_launch(n, L.out_ptr.p, L.out_idx.p, L.out_w.p, L.mod.p, L.p.p, L.tau.p, c.dangling.p, c.all_log_all, false, w,
            c.scratch, s);
  double h[8];
  RMX_HIP(hipMemcpyAsync(h, c.scal.p + 8, sizeof(h), hipMemcpyDeviceToHost, s));
  c.sync();
  return h[kL];
}

size_t p5_partial_doubles() { return 5 * (size_t)kRedGrid; }

}  // namespace rmx
