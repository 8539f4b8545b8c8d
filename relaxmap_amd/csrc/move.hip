// Map-equation move sweep (Network::prioritize_parallelMove and its twins,
// Module.cpp:962-1685 and 2154-2869) as batched relaxed moves.
//
// A sweep visits every active unit once, in B sub-sweeps ("batches") chosen
// by a counter-based hash of (seed, sweep, unit) in place of the reference's
// serial shuffle (Module.cpp:1327-1338).  Within a batch every unit is
// evaluated against the module statistics as they stand at the start of the
// batch (the reference evaluates against racy snapshots too,
// Module.cpp:1039-1041); the moves of the batch are then committed with
// exact integer updates of the module statistics, so each batch starts from
// exact statistics.  Swaps of two singletons into each other's module are
// excluded by letting a singleton join another singleton only if that
// module's id is smaller.
//
// Evaluation of one unit u in module o (Module.cpp:1361-1485):
//   for each neighbour module m: A_m = sum of flows u->m, B_m = m->u,
//   out(m) = beta A_m + (alpha p + beta d) T'_m,  in(m) = beta B_m + (alpha P'_m + beta D'_m) tau
//   (T'_o = T_o - tau, P'_o = P_o - p, D'_o = D_o - d),
//   q1 = q_o - q_u + out(o) + in(o),  q2 = q_m + q_u - out(m) - in(m),
//   dL = [phi(S + q1 + q2 - q_o - q_m) - phi(S)] - 2 [phi(q1) + phi(q2) - phi(q_o) - phi(q_m)]
//        + [phi(q1 + P_o - p) + phi(q2 + P_m + p) - phi(q_o + P_o) - phi(q_m + P_m)];
// the best candidate (smallest dL, ties to the smaller module id, dL < 0
// only) moves if dL < vThresh.  An empty module is a candidate with zero
// statistics when the unit's module has other members (Module.cpp:1435-1439).
//
// Neighbour-module flows are aggregated in an LDS hash table per unit with
// 64-bit integer atomics (exact, order independent).  Units are bucketed by
// degree: groups of 8..64 lanes per unit with LDS tables of 32..1024 keys,
// a 256-lane workgroup with a 4096-key LDS table, and for hubs a workgroup
// with its table in global memory.
#include "devutil.h"

#include <vector>

namespace rmx {

namespace {

enum { kS = 0, kL = 3, kNM = 7 };
enum { cMovers = 0, cWant = 1, cMoved = 2, cVisitsLo = 4 };
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kInvalid = 0xfffffffeu;

struct EvalArgs {
  const uint32_t *list;
  int32_t list_n;
  const uint32_t *optr, *oidx, *iptr, *iidx;
  ArcW aout, ain;
  const uint32_t *mod;
  const long long *up, *ut, *ud, *uw;
  const long long *mP, *mT, *mD, *mX;
  const int32_t *mUnits, *mStale;
  const uint8_t *act;
  int all_active;
  uint64_t seed;
  int batches, batch;
  int node_level;
  int empty_avail;
  uint32_t empty_key;
  const double *scal;
  double vthresh;
  uint32_t *target;
  long long *agg;        // per unit: flows to/from the old and the new module
  uint32_t *movers;
  uint32_t *counters;
  uint32_t *want;
  unsigned long long *visits;
  // hub tables
  uint32_t *hkeys;
  long long *hout, *hin;
  int hcap;
};

__device__ inline bool in_batch(const EvalArgs &a, uint32_t u) {
  if (!a.all_active && !a.act[u]) return false;
  if (a.batches <= 1) return true;
  return (uint32_t)(mix64(a.seed ^ ((uint64_t)u * 0x9e3779b97f4a7c15ull)) % (uint64_t)a.batches) == (uint32_t)a.batch;
}

struct Best {
  double d;
  uint32_t m;
};

__device__ inline bool better(double d, uint32_t m, const Best &b) { return d < b.d || (d == b.d && m < b.m); }

// Per-unit constants of the evaluation.
struct UnitCtx {
  uint32_t o;
  long long p, t, d;
  double pd, td, dd, qu, S;
  long long Po, To, Do, Xo;
  double qo, q1, phi_qo, phi_qoP, phi_q1, phi_q1P, phi_S;
  bool offer_empty;
};

__device__ inline double delta_for(const UnitCtx &U, long long Pm, long long Tm, long long Dm, long long Xm,
                                   long long fo, long long fi) {
  const double qm = module_exit(Pm, Tm, Dm, Xm);
  const double outn = kBeta * fx_to(fo) + (kAlpha * U.pd + kBeta * U.dd) * fx_to(Tm);
  const double inn = kBeta * fx_to(fi) + (kAlpha * fx_to(Pm) + kBeta * fx_to(Dm)) * U.td;
  const double q2 = qm + U.qu - outn - inn;
  const double Sn = U.S + U.q1 + q2 - U.qo - qm;
  const double d_all = plogp(Sn) - U.phi_S;
  const double d_exit = U.phi_q1 + plogp(q2) - U.phi_qo - plogp(qm);
  const double d_stay = U.phi_q1P + plogp(q2 + fx_to(Pm + U.p)) - U.phi_qoP - plogp(qm + fx_to(Pm));
  return d_all - 2.0 * d_exit + d_stay;
}

__device__ inline double delta_empty(const UnitCtx &U) {
  // zero statistics and no flows to the new module (Module.cpp:1220-1223)
  const double q2 = U.qu;
  const double Sn = U.S + U.q1 + q2 - U.qo;
  const double d_all = plogp(Sn) - U.phi_S;
  const double d_exit = U.phi_q1 + plogp(q2) - U.phi_qo;
  const double d_stay = U.phi_q1P + plogp(q2 + U.pd) - U.phi_qoP;
  return d_all - 2.0 * d_exit + d_stay;
}

__device__ inline void unit_ctx(const EvalArgs &a, uint32_t u, long long fo_o, long long fi_o, UnitCtx &U) {
  U.o = a.mod[u];
  U.p = a.up[u];
  U.t = a.ut[u];
  U.d = a.ud[u];
  U.pd = fx_to(U.p);
  U.td = fx_to(U.t);
  U.dd = fx_to(U.d);
  U.qu = module_exit(U.p, U.t, U.d, a.uw[u]);
  U.Po = a.mP[U.o];
  U.To = a.mT[U.o];
  U.Do = a.mD[U.o];
  U.Xo = a.mX[U.o];
  U.qo = module_exit(U.Po, U.To, U.Do, U.Xo);
  const double outo = kBeta * fx_to(fo_o) + (kAlpha * U.pd + kBeta * U.dd) * fx_to(U.To - U.t);
  const double ino = kBeta * fx_to(fi_o) + (kAlpha * fx_to(U.Po - U.p) + kBeta * fx_to(U.Do - U.d)) * U.td;
  U.q1 = U.qo - U.qu + outo + ino;
  U.S = a.scal[kS];
  U.phi_S = plogp(U.S);
  U.phi_qo = plogp(U.qo);
  U.phi_qoP = plogp(U.qo + fx_to(U.Po));
  U.phi_q1 = plogp(U.q1);
  U.phi_q1P = plogp(U.q1 + fx_to(U.Po - U.p));
  U.offer_empty = a.empty_avail && (a.node_level ? a.mStale[U.o] > 1 : a.mUnits[U.o] > 1);
}

__device__ inline bool allowed(const EvalArgs &a, const UnitCtx &U, uint32_t m) {
  // a singleton may join another singleton only if that module's id is smaller
  return !(m > U.o && a.mUnits[U.o] == 1 && a.mUnits[m] == 1);
}

__device__ inline void emit(const EvalArgs &a, uint32_t u, const Best &b, long long fo_o, long long fi_o,
                            long long fo_n, long long fi_n) {
  if (!(b.d < a.vthresh) || b.m == kNone) return;
  a.target[u] = b.m;
  long long *g = a.agg + 4 * (int64_t)u;
  g[0] = fo_o;
  g[1] = fi_o;
  g[2] = fo_n;
  g[3] = fi_n;
  uint32_t i = atomicAdd(&a.counters[cMovers], 1u);
  a.movers[i] = u;
  if (b.m == a.empty_key) {
    a.want[u] = 1u;
    atomicAdd(&a.counters[cWant], 1u);
  }
}

// ------------------------------------------------------------------ groups
template <int G>
__device__ inline Best group_best(Best b) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    double d = __shfl_xor(b.d, off, G);
    uint32_t m = __shfl_xor(b.m, off, G);
    if (better(d, m, b)) { b.d = d; b.m = m; }
  }
  return b;
}

template <int G>
__device__ inline long long group_sum_ll(long long v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// Open-addressing table (linear probing); keys only ever go from empty to
// a module id.  The global-memory variant (hub units) reads through L2.
template <bool GLOBAL = false>
__device__ inline uint32_t ld_key(const uint32_t *keys, uint32_t h) {
  if (GLOBAL) return __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return keys[h];
}

template <bool GLOBAL = false>
__device__ inline uint32_t lds_insert(uint32_t *keys, int H, uint32_t m) {
  uint32_t h = hash_u32(m) & (uint32_t)(H - 1);
  while (true) {
    uint32_t k = ld_key<GLOBAL>(keys, h);
    if (k == m) return h;
    if (k == kEmptyKey) {
      uint32_t old = atomicCAS(&keys[h], kEmptyKey, m);
      if (old == kEmptyKey || old == m) return h;
    }
    h = (h + 1) & (uint32_t)(H - 1);
  }
}

template <bool GLOBAL = false>
__device__ inline int lds_find(const uint32_t *keys, int H, uint32_t m) {
  uint32_t h = hash_u32(m) & (uint32_t)(H - 1);
  while (true) {
    uint32_t k = ld_key<GLOBAL>(keys, h);
    if (k == m) return (int)h;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & (uint32_t)(H - 1);
  }
}

// G lanes per unit (G <= 64: the group lives inside one wavefront), H keys.
template <int G, int H>
__global__ __launch_bounds__(256) void k_eval_group(EvalArgs a) {
  constexpr int GPB = 256 / G;
  __shared__ uint32_t s_keys[GPB][H];
  __shared__ long long s_out[GPB][H];
  __shared__ long long s_in[GPB][H];
  __shared__ unsigned long long s_vis[256 / kWave];
  const int gl = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int64_t g = (int64_t)blockIdx.x * GPB + gl;
  uint32_t *keys = s_keys[gl];
  long long *fo = s_out[gl];
  long long *fi = s_in[gl];
  unsigned long long visits = 0;
  bool work = false;
  uint32_t u = 0;
  if (g < a.list_n) {
    u = a.list[g];
    work = in_batch(a, u);
  }
  if (work) {
    for (int s = lane; s < H; s += G) {
      keys[s] = kEmptyKey;
      fo[s] = 0;
      fi[s] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t ob = a.optr[u], oe = a.optr[u + 1];
    const uint32_t ib = a.iptr[u], ie = a.iptr[u + 1];
    for (uint32_t k = ob + lane; k < oe; k += G) {
      uint32_t m = a.mod[a.oidx[k]];
      uint32_t s = lds_insert(keys, H, m);
      atomicAdd((unsigned long long *)&fo[s], (unsigned long long)a.aout.fxw(k));
    }
    for (uint32_t k = ib + lane; k < ie; k += G) {
      uint32_t m = a.mod[a.iidx[k]];
      uint32_t s = lds_insert(keys, H, m);
      atomicAdd((unsigned long long *)&fi[s], (unsigned long long)a.ain.fxw(k));
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) visits = (oe - ob) + (ie - ib);
    UnitCtx U;
    const uint32_t o = a.mod[u];
    const int so = lds_find(keys, H, o);
    const long long fo_o = so >= 0 ? fo[so] : 0, fi_o = so >= 0 ? fi[so] : 0;
    unit_ctx(a, u, fo_o, fi_o, U);
    Best b{0.0, kNone};
    for (int s = lane; s < H; s += G) {
      const uint32_t m = keys[s];
      if (m == kEmptyKey || m == o) continue;
      if (!allowed(a, U, m)) continue;
      const double dl = delta_for(U, a.mP[m], a.mT[m], a.mD[m], a.mX[m], fo[s], fi[s]);
      if (better(dl, m, b) && dl < 0.0) { b.d = dl; b.m = m; }
    }
    if (lane == 0 && U.offer_empty) {
      const double dl = delta_empty(U);
      if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; }
    }
    b = group_best<G>(b);
    if (lane == 0) {
      long long fo_n = 0, fi_n = 0;
      if (b.m != kNone && b.m != a.empty_key) {
        const int sn = lds_find(keys, H, b.m);
        fo_n = fo[sn];
        fi_n = fi[sn];
      }
      emit(a, u, b, fo_o, fi_o, fo_n, fi_n);
    }
  }
  // edge-visit count, one atomic per block
  visits = wave_sum(visits);
  if ((threadIdx.x & (kWave - 1)) == 0) s_vis[threadIdx.x / kWave] = visits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < 256 / kWave; ++i) t += s_vis[i];
    if (t) atomicAdd(a.visits, t);
  }
}

// One 256-lane workgroup per unit; the hash table lives in LDS (H keys) or,
// for hubs (GLOBAL), in this block's slice of a global scratch table.
template <int H, bool GLOBAL>
__global__ __launch_bounds__(256) void k_eval_block(EvalArgs a) {
  constexpr int LH = GLOBAL ? 1 : H;
  __shared__ uint32_t s_keys[LH];
  __shared__ long long s_out[LH];
  __shared__ long long s_in[LH];
  __shared__ double s_bd[256 / kWave];
  __shared__ uint32_t s_bm[256 / kWave];
  __shared__ unsigned long long s_vis;
  const int cap = GLOBAL ? a.hcap : H;
  uint32_t *keys = GLOBAL ? a.hkeys + (int64_t)blockIdx.x * cap : s_keys;
  long long *fo = GLOBAL ? a.hout + (int64_t)blockIdx.x * cap : s_out;
  long long *fi = GLOBAL ? a.hin + (int64_t)blockIdx.x * cap : s_in;
  if (threadIdx.x == 0) s_vis = 0;
  for (int64_t g = blockIdx.x; g < a.list_n; g += gridDim.x) {
    const uint32_t u = a.list[g];
    if (!in_batch(a, u)) continue;   // uniform across the block
    for (int s = threadIdx.x; s < cap; s += 256) {
      keys[s] = kEmptyKey;
      fo[s] = 0;
      fi[s] = 0;
    }
    if (GLOBAL) __threadfence();
    __syncthreads();
    const uint32_t ob = a.optr[u], oe = a.optr[u + 1];
    const uint32_t ib = a.iptr[u], ie = a.iptr[u + 1];
    for (uint32_t k = ob + threadIdx.x; k < oe; k += 256) {
      uint32_t m = a.mod[a.oidx[k]];
      uint32_t s = lds_insert<GLOBAL>(keys, cap, m);
      atomicAdd((unsigned long long *)&fo[s], (unsigned long long)a.aout.fxw(k));
    }
    for (uint32_t k = ib + threadIdx.x; k < ie; k += 256) {
      uint32_t m = a.mod[a.iidx[k]];
      uint32_t s = lds_insert<GLOBAL>(keys, cap, m);
      atomicAdd((unsigned long long *)&fi[s], (unsigned long long)a.ain.fxw(k));
    }
    if (GLOBAL) __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_vis += (oe - ob) + (ie - ib);
    UnitCtx U;
    const uint32_t o = a.mod[u];
    long long fo_o = 0, fi_o = 0;
    {
      int so = lds_find<GLOBAL>(keys, cap, o);
      if (so >= 0) {
        fo_o = GLOBAL ? __hip_atomic_load(&fo[so], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fo[so];
        fi_o = GLOBAL ? __hip_atomic_load(&fi[so], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fi[so];
      }
    }
    unit_ctx(a, u, fo_o, fi_o, U);
    Best b{0.0, kNone};
    for (int s = threadIdx.x; s < cap; s += 256) {
      const uint32_t m = GLOBAL ? __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : keys[s];
      if (m == kEmptyKey || m == o) continue;
      if (!allowed(a, U, m)) continue;
      const long long vo = GLOBAL ? __hip_atomic_load(&fo[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fo[s];
      const long long vi = GLOBAL ? __hip_atomic_load(&fi[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fi[s];
      const double dl = delta_for(U, a.mP[m], a.mT[m], a.mD[m], a.mX[m], vo, vi);
      if (better(dl, m, b) && dl < 0.0) { b.d = dl; b.m = m; }
    }
    if (threadIdx.x == 0 && U.offer_empty) {
      const double dl = delta_empty(U);
      if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; }
    }
    b = group_best<kWave>(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      s_bd[threadIdx.x / kWave] = b.d;
      s_bm[threadIdx.x / kWave] = b.m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Best bb{s_bd[0], s_bm[0]};
      for (int i = 1; i < 256 / kWave; ++i)
        if (better(s_bd[i], s_bm[i], bb)) { bb.d = s_bd[i]; bb.m = s_bm[i]; }
      long long fo_n = 0, fi_n = 0;
      if (bb.m != kNone && bb.m != a.empty_key) {
        const int sn = lds_find<GLOBAL>(keys, cap, bb.m);
        fo_n = GLOBAL ? __hip_atomic_load(&fo[sn], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fo[sn];
        fi_n = GLOBAL ? __hip_atomic_load(&fi[sn], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : fi[sn];
      }
      emit(a, u, bb, fo_o, fi_o, fo_n, fi_n);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && s_vis) atomicAdd(a.visits, s_vis);
}

// ------------------------------------------------------------------ commit
// Final target of every mover: a neighbour module, an empty module by rank
// among this batch's empty requests (ascending unit id), or invalid when no
// empty module is left (Module.cpp:1138-1152).
__global__ void k_resolve(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                          const uint32_t *__restrict__ target, const uint32_t *__restrict__ rank,
                          const uint32_t *__restrict__ empty_list, uint32_t empty_cursor, uint32_t n_empty,
                          uint32_t empty_key, uint32_t *__restrict__ newmod, uint32_t *__restrict__ stamp,
                          uint32_t *__restrict__ want, uint32_t cur) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    uint32_t t = target[u];
    if (t == empty_key) {
      const uint32_t r = empty_cursor + rank[u];
      t = r < n_empty ? empty_list[r] : kInvalid;
      want[u] = 0u;
    }
    newmod[u] = t;
    if (t != kInvalid) stamp[u] = cur;
  }
}

// G lanes per mover: module statistics, exit flows and activity marks.
// X changes of a move of u from o to n with every neighbour fixed:
//   X_o += -(W_u - A_o) + B_o,  X_n += (W_u - A_n) - B_n;
// arcs u->v whose head moves in the same batch are corrected from the tail:
//   X_o += w ([o != o_v] - [o != n_v]),  X_n += w ([n != n_v] - [n != o_v]).
template <int G>
__global__ __launch_bounds__(256) void k_commit(const uint32_t *__restrict__ movers,
                                                const uint32_t *__restrict__ counters, const uint32_t *__restrict__ optr,
                                                const uint32_t *__restrict__ oidx, const uint32_t *__restrict__ iptr,
                                                const uint32_t *__restrict__ iidx, ArcW aout,
                                                const uint32_t *__restrict__ mod, const uint32_t *__restrict__ newmod,
                                                const uint32_t *__restrict__ stamp, uint32_t cur,
                                                const long long *__restrict__ agg, const long long *__restrict__ up,
                                                const long long *__restrict__ ut, const long long *__restrict__ ud,
                                                const long long *__restrict__ uw, const int32_t *__restrict__ nmem,
                                                long long *mP, long long *mT, long long *mD, long long *mX,
                                                int32_t *mNodes, int32_t *mUnits, uint8_t *act_next, int mark,
                                                uint32_t *moved) {
  const uint32_t nm = counters[cMovers];
  const int lane = threadIdx.x % G;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x / G;
  for (int64_t i = g0; i < nm; i += gstride) {
    const uint32_t u = movers[i];
    const uint32_t ob = optr[u], oe = optr[u + 1];
    if (mark) {
      for (uint32_t k = ob + lane; k < oe; k += G) act_next[oidx[k]] = 1;
      for (uint32_t k = iptr[u] + lane; k < iptr[u + 1]; k += G) act_next[iidx[k]] = 1;
    }
    if (stamp[u] != cur) continue;
    const uint32_t o = mod[u], n = newmod[u];
    long long dxo = 0, dxn = 0;
    for (uint32_t k = ob + lane; k < oe; k += G) {
      const uint32_t v = oidx[k];
      if (stamp[v] != cur) continue;
      const long long w = aout.fxw(k);
      const uint32_t ov = mod[v], nv = newmod[v];
      dxo += w * ((long long)(o != ov) - (long long)(o != nv));
      dxn += w * ((long long)(n != nv) - (long long)(n != ov));
    }
    dxo = group_sum_ll<G>(dxo);
    dxn = group_sum_ll<G>(dxn);
    if (lane == 0) {
      const long long *g = agg + 4 * (int64_t)u;
      const long long W = uw[u];
      dxo += -(W - g[0]) + g[1];
      dxn += (W - g[2]) - g[3];
      const long long p = up[u], t = ut[u], d = ud[u];
      atomicAdd((unsigned long long *)&mP[o], (unsigned long long)(-p));
      atomicAdd((unsigned long long *)&mP[n], (unsigned long long)p);
      atomicAdd((unsigned long long *)&mT[o], (unsigned long long)(-t));
      atomicAdd((unsigned long long *)&mT[n], (unsigned long long)t);
      if (d) {
        atomicAdd((unsigned long long *)&mD[o], (unsigned long long)(-d));
        atomicAdd((unsigned long long *)&mD[n], (unsigned long long)d);
      }
      if (dxo) atomicAdd((unsigned long long *)&mX[o], (unsigned long long)dxo);
      if (dxn) atomicAdd((unsigned long long *)&mX[n], (unsigned long long)dxn);
      const int32_t k = nmem[u];
      atomicSub(&mNodes[o], k);
      atomicAdd(&mNodes[n], k);
      atomicSub(&mUnits[o], 1);
      atomicAdd(&mUnits[n], 1);
      atomicAdd(moved, (uint32_t)k);
    }
  }
}

__global__ void k_apply(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                        const uint32_t *__restrict__ newmod, const uint32_t *__restrict__ stamp, uint32_t cur,
                        uint32_t *__restrict__ mod) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    if (stamp[u] == cur) mod[u] = newmod[u];
  }
}

__global__ void k_reset_counters(uint32_t *counters) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0;
}

__global__ void k_empty_flags(const int32_t *__restrict__ mNodes, int32_t n, uint32_t *__restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = mNodes[i] == 0 ? 1u : 0u;
}

__global__ void k_compact(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, int32_t n,
                          uint32_t *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = (uint32_t)i;
}

unsigned stride_grid(int64_t items, int per_block) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

void ensure_sweep_buffers(Ctx &c) {
  const int32_t n = c.N;
  if (c.target.n >= n) return;
  c.act.alloc(n);
  c.act_next.alloc(n);
  c.target.alloc(n);
  c.newmod.alloc(n);
  c.stamp.alloc(n);
  c.movers.alloc(n);
  c.empty_list.alloc(n);
  c.want_empty.alloc(n);
  c.rank_tmp.alloc(n);
  c.counters.alloc(16);
  c.agg_buf.alloc(4 * (int64_t)n);
  fill_u8(c.act.p, 1, n, c.stream);
  fill_u8(c.act_next.p, 0, n, c.stream);
  fill_u32(c.stamp.p, 0u, n, c.stream);
  fill_u32(c.want_empty.p, 0u, n, c.stream);
  fill_u32(c.counters.p, 0u, 16, c.stream);
}

void reset_active(Ctx &c) {
  ensure_sweep_buffers(c);
  fill_u8(c.act.p, 1, c.N, c.stream);
  fill_u8(c.act_next.p, 0, c.N, c.stream);
}

int64_t move_sweep(Ctx &c, bool prioritized, double vthresh, uint64_t seed, int64_t *edge_visits) {
  ensure_sweep_buffers(c);
  Level &L = c.cur();
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  const uint64_t sweep_seed = mix64(seed ^ mix64(++c.sweep_counter));

  // empty modules available to this sweep, ascending id
  uint32_t n_empty = 0;
  {
    k_empty_flags<<<grid_for(c.N, 256), 256, 0, s>>>(c.mNodes.p, c.N, c.rank_tmp.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(c.rank_tmp.p, c.target.p, c.N, c.counters.p + 8, c.scratch, s);
    k_compact<<<grid_for(c.N, 256), 256, 0, s>>>(c.rank_tmp.p, c.target.p, c.N, c.empty_list.p);
    RMX_LAUNCH_CHECK();
  }
  fill_u32(c.counters.p + 2, 0u, 6, s);   // moved, visits
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  c.sync();
  n_empty = c.h_cnt[8];
  uint32_t empty_cursor = 0;

  // hub scratch tables
  int hcap = 0;
  const int32_t n_hubs = L.bucket_off[kNumBuckets] - L.bucket_off[kNumBuckets - 1];
  const int hub_blocks = n_hubs > 0 ? (n_hubs < 64 ? n_hubs : 64) : 0;
  if (n_hubs > 0) {
    hcap = 1;
    while (hcap < 2 * L.max_deg + 2) hcap <<= 1;
    const int64_t need = (int64_t)hub_blocks * hcap;
    if (c.hub_keys.n < need) {
      c.hub_keys.alloc(need);
      c.hub_out.alloc(need);
      c.hub_in.alloc(need);
    }
  }

  EvalArgs a;
  a.optr = L.out_ptr.p;
  a.oidx = L.out_idx.p;
  a.iptr = L.in_ptr.p;
  a.iidx = L.in_idx.p;
  a.aout = arcw_out(c, L);
  a.ain = arcw_in(c, L);
  a.mod = L.mod.p;
  a.up = L.up.p;
  a.ut = L.ut.p;
  a.ud = L.ud.p;
  a.uw = L.uw.p;
  a.mP = c.mP.p;
  a.mT = c.mT.p;
  a.mD = c.mD.p;
  a.mX = c.mX.p;
  a.mUnits = c.mUnits.p;
  a.mStale = c.mStale.p;
  a.act = c.act.p;
  a.all_active = prioritized ? 0 : 1;
  a.seed = sweep_seed;
  a.batches = c.batches;
  a.node_level = c.level == 0 ? 1 : 0;
  a.empty_key = (uint32_t)c.N + 1u;
  a.scal = c.scal.p;
  a.vthresh = prioritized ? vthresh : 0.0;
  a.target = c.target.p;
  a.agg = c.agg_buf.p;
  a.movers = c.movers.p;
  a.counters = c.counters.p;
  a.want = c.want_empty.p;
  a.visits = (unsigned long long *)(c.counters.p + cVisitsLo);
  a.hkeys = c.hub_keys.p;
  a.hout = c.hub_out.p;
  a.hin = c.hub_in.p;
  a.hcap = hcap;

  ArcW aout = arcw_out(c, L);
  for (int b = 0; b < c.batches; ++b) {
    a.batch = b;
    a.empty_avail = empty_cursor < n_empty ? 1 : 0;
    k_reset_counters<<<1, 64, 0, s>>>(c.counters.p);
    RMX_LAUNCH_CHECK();
    for (int bk = 0; bk < kNumBuckets; ++bk) {
      const int32_t cnt = L.bucket_off[bk + 1] - L.bucket_off[bk];
      if (cnt <= 0) continue;
      a.list = L.bucket.p + L.bucket_off[bk];
      a.list_n = cnt;
      switch (bk) {
        case 0: k_eval_group<8, 32><<<(unsigned)((cnt + 31) / 32), 256, 0, s>>>(a); break;
        case 1: k_eval_group<16, 64><<<(unsigned)((cnt + 15) / 16), 256, 0, s>>>(a); break;
        case 2: k_eval_group<32, 256><<<(unsigned)((cnt + 7) / 8), 256, 0, s>>>(a); break;
        case 3: k_eval_group<64, 1024><<<(unsigned)((cnt + 3) / 4), 256, 0, s>>>(a); break;
        case 4: k_eval_block<4096, false><<<(unsigned)(cnt < 8192 ? cnt : 8192), 256, 0, s>>>(a); break;
        default: k_eval_block<1, true><<<(unsigned)hub_blocks, 256, 0, s>>>(a); break;
      }
      RMX_LAUNCH_CHECK();
    }
    RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, s));
    c.sync();
    const uint32_t n_movers = c.h_cnt[cMovers], n_want = c.h_cnt[cWant];
    if (n_movers == 0) continue;
    ++c.stamp_now;
    if (n_want > 0) exclusive_scan_u32(c.want_empty.p, c.rank_tmp.p, n, nullptr, c.scratch, s);
    k_resolve<<<stride_grid(n_movers, 256), 256, 0, s>>>(c.movers.p, c.counters.p, c.target.p, c.rank_tmp.p,
                                                          c.empty_list.p, empty_cursor, n_empty, a.empty_key,
                                                          c.newmod.p, c.stamp.p, c.want_empty.p, c.stamp_now);
    RMX_LAUNCH_CHECK();
    if (n_want > 0) {
      uint32_t left = n_empty - empty_cursor;
      empty_cursor += n_want < left ? n_want : left;
    }
    k_commit<16><<<stride_grid((int64_t)n_movers * 16, 256), 256, 0, s>>>(
        c.movers.p, c.counters.p, L.out_ptr.p, L.out_idx.p, L.in_ptr.p, L.in_idx.p, aout, L.mod.p, c.newmod.p,
        c.stamp.p, c.stamp_now, c.agg_buf.p, L.up.p, L.ut.p, L.ud.p, L.uw.p, L.nmem.p, c.mP.p, c.mT.p, c.mD.p,
        c.mX.p, c.mNodes.p, c.mUnits.p, c.act_next.p, prioritized ? 1 : 0, c.counters.p + cMoved);
    RMX_LAUNCH_CHECK();
    k_apply<<<stride_grid(n_movers, 256), 256, 0, s>>>(c.movers.p, c.counters.p, c.newmod.p, c.stamp.p,
                                                        c.stamp_now, L.mod.p);
    RMX_LAUNCH_CHECK();
    module_totals_async(c);   // S for the next batch
  }
  if (prioritized) {
    std::swap(c.act.p, c.act_next.p);
    fill_u8(c.act_next.p, 0, c.N, s);
  }
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  module_totals(c, &c.sum_exit, &c.codelength);
  const uint32_t moved = c.h_cnt[cMoved];
  unsigned long long vis = (unsigned long long)c.h_cnt[cVisitsLo] | ((unsigned long long)c.h_cnt[cVisitsLo + 1] << 32);
  if (edge_visits) *edge_visits = (int64_t)vis;
  c.timing.edge_visits += (int64_t)vis;
  c.timing.sweeps++;
  return moved;
}

int64_t n_active(Ctx &c) {
  std::vector<uint8_t> h(c.N);
  RMX_HIP(hipMemcpyAsync(h.data(), c.act.p, c.N, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  int64_t k = 0;
  for (int32_t i = 0; i < c.cur().n; ++i) k += h[i];
  return k;
}

}  // namespace rmx
