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
// a 256-lane workgroup with a 2048- or 4096-key LDS table, and hubs split by
// neighbour module into partitions of about kHubPart arcs, each aggregated
// in an LDS table (hubs whose partition overflows fall back to global-memory
// tables).
#include "devutil.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace rmx {

namespace {

enum { kS = 0, kL = 3, kNM = 7 };
enum { cMovers = 0, cWant = 1, cMoved = 2, cVisitsLo = 4, cCandsLo = 6, cUnits = 10, cBig = 11, cHubOvf = 12,
       cErr = 13, cPull = 14 };

// Counters that every workgroup of a large launch updates (movers, empty
// requests, moved nodes, edge visits, candidates, units) are kept in
// kStripes copies 128 bytes apart; a workgroup uses copy (block id mod
// kStripes).  Device-scope atomics on one address serialise at about 11 ns
// each on MI355X, so a single counter bumped by every workgroup of an
// 85k-workgroup launch costs about a millisecond per update per workgroup;
// spread over the copies they proceed in parallel.  Each copy appends its
// movers to its own region of movers_st (mv_cap entries: at most
// n/64 + 1024 movers of one batch fall on one copy, since every launch maps
// at most ceil(grid/kStripes) workgroups of bounded unit count to a copy).
// k_gather_movers appends the copies' movers to the dense mover list and
// k_fold_stripes adds the copies to `counters` and clears them.
constexpr int kStripes = 128;
constexpr int kStripeWords = 32;
constexpr int kStripeBase = 512;   // offset of the copies in the counter buffer
constexpr int kCounterWords = kStripeBase + kStripes * kStripeWords;

__device__ inline uint32_t *my_stripe(uint32_t *stripes) {
  return stripes + (blockIdx.x & (uint32_t)(kStripes - 1)) * (uint32_t)kStripeWords;
}
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kInvalid = 0xfffffffeu;

struct EvalArgs {
  const uint32_t *list;     // positions into `bucket` of this launch's units
  int32_t list_n;
  const uint32_t *bucket;   // units in degree-bucket order
  int32_t hub_base;         // position of the first hub in `bucket`
  const uint32_t *optr, *oidx, *iptr, *iidx;
  ArcW aout, ain;
  const uint32_t *mod;
  const Stat4 *us;
  const Stat4 *ms;
  const uint32_t *mOne;     // bit m: module m holds one unit
  const uint32_t *mMulti;   // node level, bit m: module m held more than one node as the sweep started
  const uint8_t *act;
  int all_active;
  uint64_t seed;
  int batches, batch;
  int node_level;
  int empty_avail;
  uint32_t empty_key;
  const double *scal;
  const double *S_inst;      // per-instance S (sub-module partition or batch), indexed by inst[u]; null: scal[kS]
  const uint32_t *inst;
  const uint32_t *e_cnt, *e_cur;   // batch: empty modules of each instance and those taken (else null)
  double vthresh;
  uint32_t *target;
  long long *agg;        // per unit: flows to/from the old and the new module
  uint32_t *movers_st;   // movers per counter copy, mv_cap entries each
  uint32_t mv_cap;
  uint32_t *counters;
  uint32_t *stripes;     // kStripes copies of the per-launch counters
  uint32_t *want;        // empty-module requests of the batch: one bit per unit
  // hub tables: region [hub_off[g], hub_off[g+1]) of the pool for hub g
  uint32_t *hkeys;
  long long *hout, *hin;
  const long long *hub_off;
  // evaluation only (rmx_evaluate_moves): every evaluated unit's best
  // candidate and its dL, written before the mover test (else null)
  uint32_t *best_m;
  double *best_d;
  // candidate bound: smax[0] / smax[1] are upper bounds of every module's
  // teleport weight T and flow P (fixed point) at the batch start; bound is
  // on (1) or off (0) for this launch
  const unsigned long long *smax;
  int bound;
};

// Every unit draws its batch independently.  (Drawing by blocks of
// consecutive ids would coalesce the per-unit reads, but on networks whose
// communities occupy id ranges a whole community then moves in one batch
// against stale statistics, which was measured to raise the codelength.)
__device__ inline uint32_t batch_of(uint64_t seed, int batches, uint32_t u) {
  if (batches <= 1) return 0;
  return (uint32_t)(mix64(seed ^ ((uint64_t)u * 0x9e3779b97f4a7c15ull)) % (uint64_t)batches);
}

// module m holds exactly one unit
__device__ inline bool one_unit(const EvalArgs &a, uint32_t m) { return (a.mOne[m >> 5] >> (m & 31u)) & 1u; }

__device__ inline void put_mover(const EvalArgs &a, uint32_t slot, uint32_t u) {
  if (slot < a.mv_cap) a.movers_st[(size_t)(blockIdx.x & (uint32_t)(kStripes - 1)) * a.mv_cap + slot] = u;
  else atomicOr(&a.counters[cErr], 1u);   // cannot happen within the bound above; reported as an error
}

__device__ inline void add_visits(const EvalArgs &a, unsigned long long v) {
  atomicAdd((unsigned long long *)(my_stripe(a.stripes) + cVisitsLo), v);
}
__device__ inline void add_cands(const EvalArgs &a, unsigned long long v) {
  atomicAdd((unsigned long long *)(my_stripe(a.stripes) + cCandsLo), v);
}
__device__ inline void add_units(const EvalArgs &a, uint32_t v) { atomicAdd(my_stripe(a.stripes) + cUnits, v); }

struct Best {
  double d;
  uint32_t m;
  int i;   // slot of m in the (compacted) table, -1 for none or the empty module
};

__device__ inline bool better(double d, uint32_t m, const Best &b) { return d < b.d || (d == b.d && m < b.m); }

// Per-unit constants of the evaluation.  With d1 = q1 - q_o and
// Kr = -2 phi(q1) + 2 phi(q_o) + phi(q1 + P_o - p) - phi(q_o + P_o)
// (the unit's own module), a candidate m with q2 = q_m + d2 costs
//   dL = phi(S + d1 + d2) - phi(S) - 2 phi(q2) + 2 phi(q_m) + phi(q2 + P_m + p) - phi(q_m + P_m) + Kr.
struct UnitCtx {
  uint32_t o;
  long long p;
  double tele_out;    // alpha p + beta d
  double td, qu;
  double S, d1, Kr, phiS;
  bool single, offer_empty;
};

// d2 = q2 - q_m of a move into module M with flows fo (to) and fi (from) it
__device__ inline double cand_d2(const UnitCtx &U, const Stat4 &M, long long fo, long long fi) {
  const double flows = kBeta * (fx_to(fo) + fx_to(fi));
  const double tele = U.tele_out * fx_to(M.t) + (kAlpha * fx_to(M.p) + kBeta * fx_to(M.d)) * U.td;
  return U.qu - flows - tele;
}

__device__ inline double delta_for(const UnitCtx &U, const Stat4 &M, long long fo, long long fi) {
  const double qm = module_exit(M.p, M.t, M.d, M.x);
  const double d2 = cand_d2(U, M, fo, fi);
  const double q2 = qm + d2;
  return plogp(U.S + U.d1 + d2) - U.phiS - 2.0 * plogp(q2) + 2.0 * plogp(qm) + plogp(q2 + fx_to(M.p + U.p)) -
         plogp(qm + fx_to(M.p)) + U.Kr;
}

// Mixed precision evaluates the candidates in FP32, in difference form so
// that no large terms cancel:
//   dL = dphi(S, d1 + d2) - 2 dphi(q_m, d2) + dphi(q_m + P_m, d2 + p) + Kr,
//   dphi(x, d) = phi(x + d) - phi(x) = d ln(x + d) + x log1p(d / x),
// with d1 = q1 - q_o, d2 = q2 - q_m (computed in FP64 from the exact
// statistics) and Kr the unit's own-module terms.  Only the choice of the
// move is approximate: commits and the tracked codelength stay exact.
__device__ inline float dphi32(float x, float d) {
  const float y = x + d;
  if (!(x > 0.f)) return y > 0.f ? y * __logf(y) : 0.f;
  if (!(y > 0.f)) return -x * __logf(x);
  return d * __logf(y) + x * log1pf(d / x);
}

__device__ inline double delta_for32(const UnitCtx &U, const Stat4 &M, long long fo, long long fi) {
  const double qm = module_exit(M.p, M.t, M.d, M.x);
  const float d2 = (float)cand_d2(U, M, fo, fi);
  const float fq = (float)qm;
  const float t = dphi32((float)U.S, (float)U.d1 + d2) - 2.f * dphi32(fq, d2) +
                  dphi32(fq + (float)fx_to(M.p), d2 + (float)fx_to(U.p));
  return (double)t + U.Kr;
}

template <bool F32>
__device__ inline double cand_delta(const UnitCtx &U, const Stat4 &M, long long fo, long long fi) {
  return F32 ? delta_for32(U, M, fo, fi) : delta_for(U, M, fo, fi);
}

__device__ inline double delta_empty(const UnitCtx &U) {
  // zero statistics and no flows to the new module (Module.cpp:1220-1223): d2 = q_u
  const double q2 = U.qu;
  return plogp(U.S + U.d1 + q2) - U.phiS - 2.0 * plogp(q2) + plogp(q2 + fx_to(U.p)) + U.Kr;
}

// The unit's own reads, issued ahead of its arcs so that their latency
// overlaps the arc gathers.
struct UnitLoads {
  uint32_t o;
  Stat4 su, so;
  bool single_o;      // the unit's module holds it alone
  bool multi_o;       // node level: the module held more than one node when the sweep started
  double S;
  bool empty_left;   // an empty module is still available to the unit's instance
};

__device__ inline UnitLoads load_unit(const EvalArgs &a, uint32_t u) {
  UnitLoads l;
  l.o = a.mod[u];
  l.su = a.us[u];
  const uint32_t I = a.inst ? a.inst[u] : 0u;
  l.S = a.S_inst ? a.S_inst[I] : a.scal[kS];
  l.empty_left = a.e_cnt ? a.e_cur[I] < a.e_cnt[I] : a.empty_avail != 0;
  l.so = a.ms[l.o];
  l.single_o = one_unit(a, l.o);
  l.multi_o = a.node_level && ((a.mMulti[l.o >> 5] >> (l.o & 31u)) & 1u);
  return l;
}

__device__ inline void unit_ctx(const EvalArgs &a, const UnitLoads &l, long long fo_o, long long fi_o,
                                UnitCtx &U) {
  U.o = l.o;
  const Stat4 su = l.su;
  U.p = su.p;
  const double pd = fx_to(su.p), dd = fx_to(su.d);
  U.td = fx_to(su.t);
  U.tele_out = kAlpha * pd + kBeta * dd;
  U.qu = module_exit(su.p, su.t, su.d, su.x);
  const Stat4 so = l.so;
  const double qo = module_exit(so.p, so.t, so.d, so.x);
  const double outo = kBeta * fx_to(fo_o) + U.tele_out * fx_to(so.t - su.t);
  const double ino = kBeta * fx_to(fi_o) + (kAlpha * fx_to(so.p - su.p) + kBeta * fx_to(so.d - su.d)) * U.td;
  const double q1 = qo - U.qu + outo + ino;
  const double S = l.S;
  U.Kr = -2.0 * plogp(q1) + 2.0 * plogp(qo) + plogp(q1 + fx_to(so.p - su.p)) - plogp(qo + fx_to(so.p));
  U.S = S;
  U.phiS = plogp(S);
  U.d1 = q1 - qo;
  U.single = l.single_o;
  // (the module holds the unit itself, so more than one unit = not alone)
  U.offer_empty = l.empty_left && (a.node_level ? l.multi_o : !l.single_o);
}

__device__ inline void unit_ctx(const EvalArgs &a, uint32_t u, long long fo_o, long long fi_o, UnitCtx &U) {
  unit_ctx(a, load_unit(a, u), fo_o, fi_o, U);
}

__device__ inline bool allowed(const EvalArgs &a, const UnitCtx &U, uint32_t m) {
  // a singleton may join another singleton only if that module's id is smaller
  return !(m > U.o && U.single && one_unit(a, m));
}

__device__ inline void record_best(const EvalArgs &a, uint32_t u, const Best &b) {
  if (a.best_m) {
    a.best_m[u] = b.m;
    a.best_d[u] = b.d;
  }
}

__device__ inline void emit(const EvalArgs &a, uint32_t u, const Best &b, long long fo_o, long long fi_o,
                            long long fo_n, long long fi_n) {
  record_best(a, u, b);
  if (!(b.d < a.vthresh) || b.m == kNone) return;
  a.target[u] = b.m;
  long long *g = a.agg + 4 * (int64_t)u;
  g[0] = fo_o;
  g[1] = fi_o;
  g[2] = fo_n;
  g[3] = fi_n;
  uint32_t *st = my_stripe(a.stripes);
  put_mover(a, atomicAdd(&st[cMovers], 1u), u);
  if (b.m == a.empty_key) {
    atomicOr(&a.want[u >> 5], 1u << (u & 31u));
    atomicAdd(&st[cWant], 1u);
  }
}

// Lane mask of the lanes below this one in the wavefront.
__device__ inline uint64_t lanes_below() {
  const int wl = threadIdx.x & (kWave - 1);
  return wl == 0 ? 0ull : (~0ull >> (kWave - wl));
}

// Called by every lane of a wavefront (uniform control flow); the lanes with
// `active` set emit their unit as emit() does, with one atomic on the mover
// counter per wavefront instead of one per mover.
__device__ inline void emit_wave(const EvalArgs &a, bool active, uint32_t u, const Best &b, long long fo_o,
                                 long long fi_o, long long fo_n, long long fi_n) {
  if (active) record_best(a, u, b);
  const bool mv = active && b.d < a.vthresh && b.m != kNone;
  const uint64_t bal = __ballot(mv);
  if (bal == 0) return;
  const int leader = __ffsll((unsigned long long)bal) - 1;
  uint32_t *st = my_stripe(a.stripes);
  uint32_t base = 0;
  if ((int)(threadIdx.x & (kWave - 1)) == leader) base = atomicAdd(&st[cMovers], (uint32_t)__popcll(bal));
  base = __shfl(base, leader, kWave);
  if (!mv) return;
  a.target[u] = b.m;
  long long *g = a.agg + 4 * (int64_t)u;
  g[0] = fo_o;
  g[1] = fi_o;
  g[2] = fo_n;
  g[3] = fi_n;
  put_mover(a, base + (uint32_t)__popcll(bal & lanes_below()), u);
  if (b.m == a.empty_key) {
    atomicOr(&a.want[u >> 5], 1u << (u & 31u));
    atomicAdd(&st[cWant], 1u);
  }
}

// ------------------------------------------------------------------ groups
template <int G>
__device__ inline Best group_best(Best b) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    double d = __shfl_xor(b.d, off, G);
    uint32_t m = __shfl_xor(b.m, off, G);
    int i = __shfl_xor(b.i, off, G);
    if (better(d, m, b)) { b.d = d; b.m = m; b.i = i; }
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

// LDS tables try the compare-and-swap first: one round trip both when the
// key is new and when it is already there (global tables read first, since
// an atomic there costs more than a load).
template <bool GLOBAL = false>
__device__ inline uint32_t lds_insert(uint32_t *keys, int H, uint32_t m) {
  uint32_t h = hash_u32(m) & (uint32_t)(H - 1);
  if (!GLOBAL) {
    while (true) {
      const uint32_t old = atomicCAS(&keys[h], kEmptyKey, m);
      if (old == kEmptyKey || old == m) return h;
      h = (h + 1) & (uint32_t)(H - 1);
    }
  }
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

// Fixed-point weight of out-arc (out) or in-arc k.
__device__ inline long long arc_fxw(const EvalArgs &a, bool out, uint32_t k) {
  return *((out ? a.aout.wx : a.ain.wx) + k);
}

// Aggregates arcs j = first, first + STEP, ... of a unit's out-list followed
// by its in-list (ntot arcs, nout of them out-arcs) into the table.  Each
// lane loads U arcs (their heads, modules and weights) before inserting any,
// so that U gathers per lane are in flight at once.
template <int U, int STEP>
__device__ inline void gather_arcs(const EvalArgs &a, uint32_t ob, uint32_t nout, uint32_t ib, uint32_t ntot,
                                   uint32_t first, uint32_t *keys, int H, long long *fo, long long *fi) {
  for (uint32_t j0 = first; j0 < ntot; j0 += (uint32_t)(U * STEP)) {
    uint32_t v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t j = j0 + (uint32_t)(q * STEP);
      const uint32_t *p = j < nout ? a.oidx + ob + j : a.iidx + ib + (j - nout);
      v[q] = j < ntot ? *p : 0u;
    }
    uint32_t m[U];
    long long w[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t j = j0 + (uint32_t)(q * STEP);
      const bool out = j < nout;
      m[q] = j < ntot ? a.mod[v[q]] : 0u;
      w[q] = j < ntot ? arc_fxw(a, out, out ? ob + j : ib + (j - nout)) : 0ll;
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t j = j0 + (uint32_t)(q * STEP);
      if (j < ntot) {
        const uint32_t sl = lds_insert(keys, H, m[q]);
        atomicAdd((unsigned long long *)(j < nout ? &fo[sl] : &fi[sl]), (unsigned long long)w[q]);
      }
    }
  }
}

// Best candidate among table slots slot(first), slot(first + STEP), ...
// below n; UC candidates per lane have their module statistics loaded
// together.
template <int UC, int STEP, bool F32, typename Slot>
__device__ inline void eval_cands(const EvalArgs &a, const UnitCtx &U, const uint32_t *keys, const long long *fo,
                                  const long long *fi, int first, int n, Slot slot, Best &b) {
  for (int j0 = first; j0 < n; j0 += UC * STEP) {
    uint32_t mm[UC];
    int sl[UC];
#pragma unroll
    for (int q = 0; q < UC; ++q) {
      const int j = j0 + q * STEP;
      sl[q] = slot(j < n ? j : first);
      mm[q] = j < n ? keys[sl[q]] : U.o;
    }
    // a singleton may join another singleton only if that module's id is
    // smaller: the rule is decided from the cached one-unit bits before any
    // module row is read, so excluded candidates cost no memory request
    bool skip[UC];
#pragma unroll
    for (int q = 0; q < UC; ++q)
      skip[q] = mm[q] == U.o || (U.single && mm[q] > U.o && one_unit(a, mm[q]));
    Stat4 M[UC];
#pragma unroll
    for (int q = 0; q < UC; ++q) M[q] = skip[q] ? Stat4{0, 0, 0, 0} : a.ms[mm[q]];
#pragma unroll
    for (int q = 0; q < UC; ++q) {
      if (skip[q]) continue;
      const double dl = cand_delta<F32>(U, M[q], fo[sl[q]], fi[sl[q]]);
      if (better(dl, mm[q], b) && dl < 0.0) {
        b.d = dl;
        b.m = mm[q];
        b.i = sl[q];
      }
    }
  }
}

// ------------------------------------------------------- candidate bound
// A lower bound of a candidate's dL from the unit's flow f to it alone, so
// that a candidate that cannot beat a threshold is dropped before its module
// row is read.  With X = q_m, Y = q_m + P_m and d = q_u - f - tele_m,
//   dL(m) = A(d) + F(X, Y, d) + Kr,   A(d) = phi(S + d1 + d) - phi(S),
//   F = -2 [phi(X + d) - phi(X)] + [phi(Y + d + p) - phi(Y)].
// Every module's exit is at most its flow (q_m <= P_m, so Y >= 2X) and a
// unit's exit at most its flow (q_u <= p, so d <= p).  F grows with Y when
// d + p >= 0 and F(X, 2X, d) grows with X when p >= d, so over all modules
//   F >= phi(d + p) - 2 phi(d)                      (d >= 0: X = Y = 0)
//   F >= 2 phi(-d) + phi(p - d) - phi(-2d)          (d < 0: X = -d, Y = -2d).
// A + F_min grows with d: for d >= 0 its slope is ln((S + d1 + d)(d + p) / d^2)
// with S + d1 >= 0, and for d < 0 it is ln(4 (S + d1 + d) / (p - d)), which
// is checked at the lowest d.  The lowest d comes from the largest
// teleportation flow between the unit and any module,
//   tele_m <= tele_out T_max + P_max tau_u   (alpha P_m + beta D_m <= P_m).
// A candidate is dropped when bound - margin > threshold, the margin covering
// the rounding of the evaluated dL (FP32 terms in mixed precision): it then
// evaluates above the threshold, so it can be neither the best nor a move.
struct BoundCtx {
  double p, tm;
  bool on;
};

__device__ inline BoundCtx bound_ctx(const EvalArgs &a, const UnitCtx &U) {
  BoundCtx B;
  B.p = fx_to(U.p);
  const double tmax = fx_to((long long)a.smax[0]), pmax = fx_to((long long)a.smax[1]);
  B.tm = U.tele_out * tmax + pmax * U.td;
  B.on = U.qu <= B.p * (1.0 + 1e-12) && U.S > 0.0;
  return B;
}

// true: the candidate with flows fo + fi (fixed point) evaluates above thr.
// The bound's terms are evaluated in FP32 in difference form,
//   A = e ln(S + e) + S log1p(e / S) with e = d1 + d,
//   F_min = (d + p) ln(d + p) - 2 d ln d          (d >= 0)
//         = (p - d) ln(p - d) + 2 d ln 2          (d < 0),
// with a margin of 1e-3 of the terms' scale: far above the FP32 rounding of
// both the bound and the evaluated dL (about 1e-6 of that scale).
__device__ inline bool bounded_out(const UnitCtx &U, const BoundCtx &B, long long fsum, double thr) {
  const double dhi = U.qu - kBeta * fx_to(fsum);
  const double dlo = dhi - B.tm;
  const double p = B.p;
  if (!(dlo + p >= 0.0)) return false;
  const double e = U.d1 + dlo;
  if (!(dlo >= 0.0) && !(4.0 * (U.S + e) >= p - dlo)) return false;
  const float S = (float)U.S, ef = (float)e, y = (float)(U.S + e);
  const float A = y > 0.f ? ef * __logf(y) + S * log1pf(ef / S) : -S * __logf(S);
  const float d = (float)dlo, pf = (float)p;
  float F;
  if (dlo >= 0.0) F = (d + pf) * __logf(d + pf) - (d > 0.f ? 2.f * d * __logf(d) : 0.f);
  else F = (pf - d) * __logf(pf - d) + 1.3862944f * d;
  const double scale = fabs(U.d1) + 4.0 * fmax(fabs(dlo), fabs(dhi)) + p;
  return (double)(A + F) + U.Kr - (1e-3 * scale + 1e-11) > thr;
}

// eval_cands over a group of G lanes (G <= 64, inside one wavefront; gmask
// its lanes).  With BOUND and a.bound set, the candidates whose bound lies
// above dL = 0 (so that they cannot be a move), those of the unit's own
// module and those the singleton rule excludes are first dropped, the rest
// compacted in place to the front, in order (a slot only moves down and
// every lane of the group reads before any lane writes), and evaluated.
// *rows: the module rows read.  (A tighter threshold, the dL of the unit's
// largest-flow candidate evaluated first or of the first round of
// candidates, drops more rows but made the kernels slower: they wait on
// chains of dependent reads, and the first costs one more per unit.)
template <bool BOUND, int UC, int G, bool F32, typename Slot>
__device__ inline void eval_cands_bounded(const EvalArgs &a, const UnitCtx &U, uint32_t *keys, long long *fo,
                                          long long *fi, int lane, int n, uint64_t gmask, Slot slot, Best &b,
                                          int *rows) {
  int cnt = n;
  if (BOUND && a.bound != 0 && n > 1) {
    const BoundCtx B = bound_ctx(a, U);
    cnt = 0;
    for (int j0 = 0; j0 < n; j0 += G) {
      const int j = j0 + lane;
      uint32_t key = kEmptyKey;
      long long vo = 0, vi = 0;
      bool keep = false;
      if (j < n) {
        const int s = slot(j);
        key = keys[s];
        vo = fo[s];
        vi = fi[s];
        keep = !(key == U.o || (U.single && key > U.o && one_unit(a, key)));
        if (keep && B.on) keep = !bounded_out(U, B, vo + vi, 0.0);
      }
      const uint64_t bal = __ballot(keep) & gmask;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (keep) {
        const int d = slot(cnt + (int)__popcll(bal & lanes_below()));
        keys[d] = key;
        fo[d] = vo;
        fi[d] = vi;
      }
      cnt += (int)__popcll(bal);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  eval_cands<UC, G, F32>(a, U, keys, fo, fi, lane, cnt, slot, b);
  *rows = cnt;
}

// G lanes per unit (G <= 64: the group lives inside one wavefront), H keys.
// After the flows are aggregated, the non-empty slots are compacted in place
// to the front of the table (a slot only ever moves down, and every lane
// reads its slot before any lane writes), so the FP64 evaluation runs on
// full lanes instead of once per table stride with most lanes masked.
template <int G, int H, bool F32>
__global__ __launch_bounds__(256) void k_eval_group(EvalArgs a) {
  constexpr int GPB = 256 / G;
  __shared__ uint32_t s_keys[GPB][H];
  __shared__ long long s_out[GPB][H];
  __shared__ long long s_in[GPB][H];
  const int gl = threadIdx.x / G;
  const int lane = threadIdx.x % G;
  const int64_t g = (int64_t)blockIdx.x * GPB + gl;
  uint32_t *keys = s_keys[gl];
  long long *fo = s_out[gl];
  long long *fi = s_in[gl];
  unsigned long long visits = 0, ncand = 0;
  bool work = false;
  uint32_t u = 0;
  if (g < a.list_n) {
    u = a.list[g];   // the batch list holds the units themselves below the hub bucket
    work = true;
  }
  Best b{0.0, kNone, -1};
  long long fo_o = 0, fi_o = 0, fo_n = 0, fi_n = 0;
  if (work) {
    const uint32_t ob = a.optr[u], oe = a.optr[u + 1];
    const uint32_t ib = a.iptr[u], ie = a.iptr[u + 1];
    const UnitLoads ul = load_unit(a, u);
    for (int s = lane; s < H; s += G) {
      keys[s] = kEmptyKey;
      fo[s] = 0;
      fi[s] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    gather_arcs<(G >= 32 ? 4 : 2), G>(a, ob, oe - ob, ib, (oe - ob) + (ie - ib), (uint32_t)lane, keys, H, fo, fi);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) visits = (oe - ob) + (ie - ib);
    const uint32_t o = ul.o;
    const int so = lds_find(keys, H, o);
    if (so >= 0) {
      fo_o = fo[so];
      fi_o = fi[so];
    }
    // compact the non-empty slots to [0, nc) (the group's lanes of the wavefront)
    const int wl = threadIdx.x & (kWave - 1);
    const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1ull) << (wl - lane));
    int nc = 0;
    for (int s0 = 0; s0 < H; s0 += G) {
      const int s = s0 + lane;
      const uint32_t key = keys[s];
      const long long vo = fo[s], vi = fi[s];
      const bool ne = key != kEmptyKey;
      const uint64_t bal = __ballot(ne) & gmask;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (ne) {
        const int d = nc + (int)__popcll(bal & lanes_below());
        keys[d] = key;
        fo[d] = vo;
        fi[d] = vi;
      }
      nc += (int)__popcll(bal);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    UnitCtx U;
    unit_ctx(a, ul, fo_o, fi_o, U);
    int rows = 0;
    eval_cands_bounded<(G >= 32), (G >= 32 ? 2 : 1), G, F32>(a, U, keys, fo, fi, lane, nc, gmask,
                                                             [](int j) { return j; }, b, &rows);
    if (lane == 0) ncand = (unsigned long long)rows;
    if (lane == 0 && U.offer_empty) {
      const double dl = delta_empty(U);
      if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; b.i = -1; }
    }
    b = group_best<G>(b);
    if (lane == 0 && b.i >= 0) {
      fo_n = fo[b.i];
      fi_n = fi[b.i];
    }
  }
  emit_wave(a, work && lane == 0, u, b, fo_o, fi_o, fo_n, fi_n);
  // edge visits, candidates and units: one atomic each per block (the
  // partial sums reuse the first table once every group is done with it, so
  // the kernel's LDS is exactly its tables)
  constexpr int NW = 256 / kWave;
  static_assert(H >= 3 * NW, "table too small for the block partials");
  unsigned long long units = (work && lane == 0) ? 1ull : 0ull;   // a unit with no arcs is still one evaluation
  visits = wave_sum(visits);
  ncand = wave_sum(ncand);
  units = wave_sum(units);
  __syncthreads();
  unsigned long long *red = (unsigned long long *)&s_out[0][0];
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[threadIdx.x / kWave] = visits;
    red[NW + threadIdx.x / kWave] = ncand;
    red[2 * NW + threadIdx.x / kWave] = units;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0, c = 0, nu = 0;
    for (int i = 0; i < NW; ++i) {
      t += red[i];
      c += red[NW + i];
      nu += red[2 * NW + i];
    }
    if (t) add_visits(a, t);
    if (c) add_cands(a, c);
    if (nu) add_units(a, (uint32_t)nu);
  }
}

// One 256-lane workgroup per unit; the LDS hash table holds up to H keys,
// sized to the unit's degree.  Candidates are compacted in place per
// wavefront: wave w owns the slots s with (s mod 256) / 64 == w and moves
// its j-th non-empty slot to its own j-th slot (a slot only moves down within
// the wave's sequence, and the wave reads before it writes), then evaluates
// its own list.  The kernel's LDS is exactly the table.
__device__ inline int wave_slot(int j, int w) { return (j >> 6) * 256 + w * kWave + (j & (kWave - 1)); }

template <int H, bool F32>
__global__ __launch_bounds__(256) void k_eval_block(EvalArgs a) {
  constexpr int NW = 256 / kWave;
  static_assert(H >= 6 * NW, "table too small for the per-wave bests");
  __shared__ __align__(16) uint32_t s_keys[H];
  __shared__ long long s_out[H];
  __shared__ long long s_in[H];
  const int w = threadIdx.x / kWave;
  const int wl = threadIdx.x & (kWave - 1);
  unsigned long long vis_acc = 0, cand_acc = 0;   // thread 0
  for (int64_t g = blockIdx.x; g < a.list_n; g += gridDim.x) {
    const uint32_t u = a.list[g];   // the batch list holds the units themselves below the hub bucket
    const uint32_t ob = a.optr[u], oe = a.optr[u + 1];
    const uint32_t ib = a.iptr[u], ie = a.iptr[u + 1];
    const UnitLoads ul = load_unit(a, u);
    int cap = 256;
    while (cap < 2 * (int)((oe - ob) + (ie - ib)) + 2) cap <<= 1;
    if (cap > H) cap = H;
    uint32_t *keys = s_keys;
    long long *fo = s_out, *fi = s_in;
    for (int s = threadIdx.x; s < cap; s += 256) {
      keys[s] = kEmptyKey;
      fo[s] = 0;
      fi[s] = 0;
    }
    __syncthreads();
    gather_arcs<4, 256>(a, ob, oe - ob, ib, (oe - ob) + (ie - ib), threadIdx.x, keys, cap, fo, fi);
    __syncthreads();
    if (threadIdx.x == 0) {
      vis_acc += (oe - ob) + (ie - ib);
      add_units(a, 1u);
    }
    const uint32_t o = ul.o;
    long long fo_o = 0, fi_o = 0;
    {
      int so = lds_find(keys, cap, o);
      if (so >= 0) {
        fo_o = fo[so];
        fi_o = fi[so];
      }
    }
    __syncthreads();   // every lookup of o is done before the table is compacted
    int nw = 0;        // this wave's candidates
    for (int s0 = 0; s0 < cap; s0 += 256) {   // cap is a multiple of 256
      const int s = s0 + w * kWave + wl;
      const uint32_t key = keys[s];
      const long long vo = fo[s], vi = fi[s];
      const bool ne = key != kEmptyKey;
      const uint64_t bal = __ballot(ne);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (ne) {
        const int d = wave_slot(nw + (int)__popcll(bal & lanes_below()), w);
        keys[d] = key;
        fo[d] = vo;
        fi[d] = vi;
      }
      nw += (int)__popcll(bal);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    UnitCtx U;
    unit_ctx(a, ul, fo_o, fi_o, U);
    Best b{0.0, kNone, -1};
    int rows = 0;
    eval_cands_bounded<true, 2, kWave, F32>(a, U, keys, fo, fi, wl, nw, ~0ull,
                                            [w](int j) { return wave_slot(j, w); }, b, &rows);
    if (threadIdx.x == 0 && U.offer_empty) {
      const double dl = delta_empty(U);
      if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; b.i = -1; }
    }
    b = group_best<kWave>(b);
    const unsigned long long nc_w = (unsigned long long)rows;
    __syncthreads();   // keys are not read again: the per-wave bests reuse them
    double *r_bd = (double *)s_keys;                      // NW doubles
    uint32_t *r_bm = s_keys + 2 * NW;                     // NW ids
    int *r_bi = (int *)(s_keys + 3 * NW);                 // NW slots
    unsigned long long *r_nc = (unsigned long long *)(s_keys + 4 * NW);   // NW counts
    if (wl == 0) {
      r_bd[w] = b.d;
      r_bm[w] = b.m;
      r_bi[w] = b.i;
      r_nc[w] = nc_w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Best bb{r_bd[0], r_bm[0], r_bi[0]};
      cand_acc += r_nc[0];
      for (int i = 1; i < NW; ++i) {
        cand_acc += r_nc[i];
        if (better(r_bd[i], r_bm[i], bb)) { bb.d = r_bd[i]; bb.m = r_bm[i]; bb.i = r_bi[i]; }
      }
      const long long fo_n = bb.i >= 0 ? fo[bb.i] : 0, fi_n = bb.i >= 0 ? fi[bb.i] : 0;
      emit(a, u, bb, fo_o, fi_o, fo_n, fi_n);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (vis_acc) add_visits(a, vis_acc);
    if (cand_acc) add_cands(a, cand_acc);
  }
}

// ---------------------------------------------------------------- hubs
// Hub units (degree > 2048) are split over many workgroups: their arcs and
// their hash-table slots are cut into chunks of kHubChunk, and every phase
// (clear, insert, evaluate) runs one workgroup per chunk, so a single huge
// super node no longer serialises a launch; a final pass reduces each hub's
// per-chunk best candidates in a fixed order.
constexpr int kHubChunk = 4096;      // table slots per evaluation workgroup
constexpr int kHubArcChunk = 4096;   // arcs per insertion workgroup

__device__ inline int chunk_owner(const uint32_t *__restrict__ pre, int cnt, uint32_t c) {
  int lo = 0, hi = cnt;   // pre[lo] <= c < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// inclusive scan over a workgroup of BS lanes; *total = the sum of all
template <int BS = 256>
__device__ inline uint32_t block_incl_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(x, off, kWave);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) lds[w] = x;
  __syncthreads();
  uint32_t pre = 0, all = 0;
  for (int i = 0; i < BS / kWave; ++i) {
    if (i < w) pre += lds[i];
    all += lds[i];
  }
  __syncthreads();
  *total = all;
  return pre + x;
}

struct HubPlan {
  uint32_t *A;      // prefix of arc chunks per hub (cnt + 1)
  uint32_t *S;      // prefix of slot chunks per hub (cnt + 1)
  uint32_t *tot;    // [0] arc chunks, [1] slot chunks
  double *bd;       // best dL per slot chunk
  uint32_t *bm;     // best module per slot chunk
  const uint32_t *only;   // finish only the hubs flagged here (null: all)
};

__device__ inline void hub_of(const EvalArgs &a, int i, uint32_t &u, long long &base, int &cap) {
  const uint32_t pos = a.list[i];
  u = a.bucket[pos];
  const uint32_t h = pos - (uint32_t)a.hub_base;
  base = a.hub_off[h];
  cap = (int)(a.hub_off[h + 1] - base);
}

__global__ __launch_bounds__(256) void k_hub_plan(EvalArgs a, HubPlan P) {
  __shared__ uint32_t lds[256 / kWave];
  uint32_t carryA = 0, carryS = 0;
  if (threadIdx.x == 0) {
    P.A[0] = 0;
    P.S[0] = 0;
  }
  for (int base = 0; base < a.list_n; base += 256) {
    const int i = base + threadIdx.x;
    uint32_t ac = 0, sc = 0;
    if (i < a.list_n) {
      uint32_t u;
      long long off;
      int cap;
      hub_of(a, i, u, off, cap);
      const uint32_t deg = (a.optr[u + 1] - a.optr[u]) + (a.iptr[u + 1] - a.iptr[u]);
      ac = deg == 0 ? 1u : (deg + kHubArcChunk - 1) / kHubArcChunk;
      sc = ((uint32_t)cap + kHubChunk - 1) / kHubChunk;
      if (P.only && !P.only[i]) ac = sc = 0;   // not flagged: no work at all
    }
    uint32_t ta, ts;
    const uint32_t ia = block_incl_scan(ac, lds, &ta);
    const uint32_t is = block_incl_scan(sc, lds, &ts);
    if (i < a.list_n) {
      P.A[i + 1] = carryA + ia;
      P.S[i + 1] = carryS + is;
    }
    carryA += ta;
    carryS += ts;
  }
  if (threadIdx.x == 0) {
    P.tot[0] = carryA;
    P.tot[1] = carryS;
  }
}

__global__ __launch_bounds__(256) void k_hub_clear(EvalArgs a, HubPlan P) {
  const uint32_t total = P.tot[1];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const int i = chunk_owner(P.S, a.list_n, c);
    uint32_t u;
    long long base;
    int cap;
    hub_of(a, i, u, base, cap);
    const int s0 = (int)(c - P.S[i]) * kHubChunk;
    const int s1 = min(cap, s0 + kHubChunk);
    for (int s = s0 + threadIdx.x; s < s1; s += 256) {
      a.hkeys[base + s] = kEmptyKey;
      a.hout[base + s] = 0;
      a.hin[base + s] = 0;
    }
  }
}

// One workgroup per chunk of kHubArcChunk arcs of a hub, straight into the
// hub's global table.  (Aggregating a chunk in LDS first was measured slower:
// a 2048-arc chunk rarely repeats a module, and the table halves occupancy.)
__global__ __launch_bounds__(256) void k_hub_insert(EvalArgs a, HubPlan P) {
  const uint32_t total = P.tot[0];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const int i = chunk_owner(P.A, a.list_n, c);
    uint32_t u;
    long long base;
    int cap;
    hub_of(a, i, u, base, cap);
    const uint32_t ob = a.optr[u], od = a.optr[u + 1] - ob;
    const uint32_t ib = a.iptr[u], deg = od + (a.iptr[u + 1] - ib);
    const uint32_t j0 = (c - P.A[i]) * kHubArcChunk;
    const uint32_t j1 = min(deg, j0 + kHubArcChunk);
    uint32_t *keys = a.hkeys + base;
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256) {
      if (j < od) {
        const uint32_t k = ob + j;
        const uint32_t sl = lds_insert<true>(keys, cap, a.mod[a.oidx[k]]);
        atomicAdd((unsigned long long *)&a.hout[base + sl], (unsigned long long)a.aout.fxw(k));
      } else {
        const uint32_t k = ib + (j - od);
        const uint32_t sl = lds_insert<true>(keys, cap, a.mod[a.iidx[k]]);
        atomicAdd((unsigned long long *)&a.hin[base + sl], (unsigned long long)a.ain.fxw(k));
      }
    }
  }
}

__device__ inline long long ld_ll(const long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void hub_unit_ctx(const EvalArgs &a, uint32_t u, long long base, int cap, UnitCtx &U,
                                    long long &fo_o, long long &fi_o) {
  const uint32_t o = a.mod[u];
  fo_o = 0;
  fi_o = 0;
  const int so = lds_find<true>(a.hkeys + base, cap, o);
  if (so >= 0) {
    fo_o = ld_ll(&a.hout[base + so]);
    fi_o = ld_ll(&a.hin[base + so]);
  }
  unit_ctx(a, u, fo_o, fi_o, U);
}

// one workgroup per slot chunk: the chunk's non-empty slots are listed in
// LDS first, then evaluated densely
template <bool F32>
__global__ __launch_bounds__(256) void k_hub_eval(EvalArgs a, HubPlan P) {
  constexpr int NW = 256 / kWave;
  __shared__ uint16_t s_list[kHubChunk];
  __shared__ uint32_t s_wcnt[NW];
  __shared__ double s_bd[NW];
  __shared__ uint32_t s_bm[NW];
  const int w = threadIdx.x / kWave;
  const uint32_t total = P.tot[1];
  unsigned long long ncand = 0;
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const int i = chunk_owner(P.S, a.list_n, c);
    uint32_t u;
    long long base;
    int cap;
    hub_of(a, i, u, base, cap);
    UnitCtx U;
    long long fo_o, fi_o;
    hub_unit_ctx(a, u, base, cap, U, fo_o, fi_o);
    const int s0 = (int)(c - P.S[i]) * kHubChunk;
    const int s1 = min(cap, s0 + kHubChunk);
    int nc = 0;
    for (int t0 = s0; t0 < s1; t0 += 256) {
      const int sl = t0 + threadIdx.x;
      const bool ne = sl < s1 && ld_key<true>(a.hkeys + base, (uint32_t)sl) != kEmptyKey;
      const uint64_t bal = __ballot(ne);
      if ((threadIdx.x & (kWave - 1)) == 0) s_wcnt[w] = (uint32_t)__popcll(bal);
      __syncthreads();
      int off = nc, tot = 0;
      for (int q = 0; q < NW; ++q) {
        if (q < w) off += (int)s_wcnt[q];
        tot += (int)s_wcnt[q];
      }
      if (ne) s_list[off + (int)__popcll(bal & lanes_below())] = (uint16_t)(sl - s0);
      nc += tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) ncand += (unsigned long long)nc;
    Best b{0.0, kNone, -1};
    for (int q = threadIdx.x; q < nc; q += 256) {
      const int sl = s0 + (int)s_list[q];
      const uint32_t m = ld_key<true>(a.hkeys + base, (uint32_t)sl);
      if (m == U.o || !allowed(a, U, m)) continue;
      const double dl = cand_delta<F32>(U, a.ms[m], ld_ll(&a.hout[base + sl]), ld_ll(&a.hin[base + sl]));
      if (better(dl, m, b) && dl < 0.0) { b.d = dl; b.m = m; }
    }
    b = group_best<kWave>(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      s_bd[w] = b.d;
      s_bm[w] = b.m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Best bb{s_bd[0], s_bm[0], -1};
      for (int q = 1; q < NW; ++q)
        if (better(s_bd[q], s_bm[q], bb)) { bb.d = s_bd[q]; bb.m = s_bm[q]; }
      P.bd[c] = bb.d;
      P.bm[c] = bb.m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && ncand) add_cands(a, ncand);
}

// one wavefront per hub: best over its chunks (fixed order), empty option, emit
__global__ __launch_bounds__(256) void k_hub_finish(EvalArgs a, HubPlan P) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t wstride = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = w0; i < a.list_n; i += wstride) {
    if (P.only && !P.only[i]) continue;
    uint32_t u;
    long long base;
    int cap;
    hub_of(a, (int)i, u, base, cap);
    Best b{0.0, kNone, -1};
    for (uint32_t c = P.S[i] + lane; c < P.S[i + 1]; c += kWave)
      if (better(P.bd[c], P.bm[c], b)) { b.d = P.bd[c]; b.m = P.bm[c]; }
    b = group_best<kWave>(b);
    if (lane == 0) {
      UnitCtx U;
      long long fo_o, fi_o;
      hub_unit_ctx(a, u, base, cap, U, fo_o, fi_o);
      if (U.offer_empty) {
        const double dl = delta_empty(U);
        if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; }
      }
      long long fo_n = 0, fi_n = 0;
      if (b.m != kNone && b.m != a.empty_key) {
        const int sn = lds_find<true>(a.hkeys + base, cap, b.m);
        fo_n = ld_ll(&a.hout[base + sn]);
        fi_n = ld_ll(&a.hin[base + sn]);
      }
      emit(a, u, b, fo_o, fi_o, fo_n, fi_n);
      add_visits(a, (unsigned long long)((a.optr[u + 1] - a.optr[u]) + (a.iptr[u + 1] - a.iptr[u])));
      add_units(a, 1u);
    }
  }
}

// ------------------------------------------------------- hubs, partitioned
// The hubs of a batch are evaluated with no global atomic per arc: their
// arcs are split into partitions of about kHubPart arcs by a hash of the
// neighbour's module (all arcs into one module land in one partition), the
// partitions are laid out contiguously (count, scan, scatter), and one
// workgroup per partition aggregates its flows in an LDS table and evaluates
// its candidates densely; a last pass takes every hub's best over its
// partitions in a fixed order.  Should a partition ever hold more distinct
// modules than its table (far beyond the expected kHubPart), its hub is
// flagged and evaluated again by the global-table kernels above.
constexpr int kHubTable = 2 * kHubPart;   // LDS slots of a partition
constexpr int kHubHist = 4096;            // LDS bins of a chunk's histogram (more: global atomics)
constexpr uint32_t kInBit = 0x80000000u;  // entry direction: in-arc
constexpr uint32_t kOwnArc = 0xffffffffu; // an arc into the hub's own module (module ids stay below 2^31 - 1)

struct HubParts {
  uint32_t *A;      // arc-chunk prefix per hub (n_h + 1)
  uint32_t *PB;     // partition prefix per hub (n_h + 1)
  uint32_t *AB;     // arc prefix per hub (n_h + 1)
  uint32_t *tot;    // [0] arc chunks, [1] partitions, [2] arcs
  uint32_t *bcnt;   // entries per partition
  uint32_t *bcur;   // the scatter's cursor per partition
  uint32_t *boff;   // first entry of each partition (scan of the counts)
  uint32_t *mtmp;   // module of each batch hub arc (kInBit: in-arc)
  uint32_t *ekey;   // entries grouped by partition (module | kInBit)
  long long *ew;    // ... and their fixed-point flows
  long long *own;   // per hub: flows to and from its own module
  uint32_t *ovf;    // per hub: some partition overflowed
  double *bd;       // per partition: best dL, module, and its flows
  uint32_t *bm;
  long long *bfo, *bfi;
  uint32_t *hubu;   // per hub: its unit
  uint32_t *chub;   // per arc chunk: its hub
  uint32_t *phub;   // per partition: its hub
  int force_ovf;    // treat every partition as overflowed (tests the fallback)
};

__device__ inline uint32_t part_of(uint32_t m, uint32_t np) {
  return (uint32_t)(((uint64_t)hash_u32(m ^ 0x9e3779b9u) * np) >> 32);
}

__device__ inline int lds_insert_bounded(uint32_t *keys, int H, uint32_t m) {
  uint32_t h = hash_u32(m) & (uint32_t)(H - 1);
  for (int probe = 0; probe < H; ++probe) {
    const uint32_t old = atomicCAS(&keys[h], kEmptyKey, m);
    if (old == kEmptyKey || old == m) return (int)h;
    h = (h + 1) & (uint32_t)(H - 1);
  }
  return -1;   // table full
}

// One workgroup of 1024 lanes.  Lane t takes the contiguous hubs
// [t k, t k + k), k = ceil(list_n / 1024): it reads their degrees (several
// in flight) and parks them in AB, and one workgroup scan of the lanes'
// totals gives every lane the prefix at its first hub.
constexpr int kPlanBlock = 1024;
__global__ __launch_bounds__(kPlanBlock) void k_hubp_plan(EvalArgs a, HubParts H) {
  __shared__ uint32_t lds[kPlanBlock / kWave];
  const int n = a.list_n;
  const int k = (n + kPlanBlock - 1) / kPlanBlock;
  const int i0 = min(n, (int)threadIdx.x * k), i1 = min(n, i0 + k);
  uint32_t ac = 0, pc = 0, dc = 0;
  for (int b = i0; b < i1; b += 4) {
    uint32_t u[4], deg[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = b + q < i1 ? a.bucket[a.list[b + q]] : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      deg[q] = b + q < i1 ? (a.optr[u[q] + 1] - a.optr[u[q]]) + (a.iptr[u[q] + 1] - a.iptr[u[q]]) : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (b + q >= i1) break;
      H.hubu[b + q] = u[q];
      H.AB[b + q + 1] = deg[q];   // this lane reads it back below
      ac += (deg[q] + kHubArcChunk - 1) / kHubArcChunk;
      pc += (deg[q] + kHubPart - 1) / kHubPart;
      dc += deg[q];
    }
  }
  uint32_t ta, tp, tb;
  uint32_t ca = block_incl_scan<kPlanBlock>(ac, lds, &ta) - ac;
  uint32_t cp = block_incl_scan<kPlanBlock>(pc, lds, &tp) - pc;
  uint32_t cb = block_incl_scan<kPlanBlock>(dc, lds, &tb) - dc;
  if (threadIdx.x == 0) {
    H.A[0] = 0;
    H.PB[0] = 0;
    H.AB[0] = 0;
    H.tot[0] = ta;
    H.tot[1] = tp;
    H.tot[2] = tb;
  }
  for (int i = i0; i < i1; ++i) {
    const uint32_t deg = H.AB[i + 1];
    ca += (deg + kHubArcChunk - 1) / kHubArcChunk;
    cp += (deg + kHubPart - 1) / kHubPart;
    cb += deg;
    H.A[i + 1] = ca;
    H.PB[i + 1] = cp;
    H.AB[i + 1] = cb;
  }
}

// the hub of every arc chunk and of every partition (a wavefront per hub),
// so that chunk and partition workgroups find their hub with one read
__global__ __launch_bounds__(256) void k_hubp_owners(EvalArgs a, HubParts H) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= a.list_n) return;
  for (uint32_t c = H.A[i] + lane; c < H.A[i + 1]; c += kWave) H.chub[c] = (uint32_t)i;
  for (uint32_t p = H.PB[i] + lane; p < H.PB[i + 1]; p += kWave) {
    H.phub[p] = (uint32_t)i;
    H.bcnt[p] = 0;   // the hub's partitions start empty
    H.bcur[p] = 0;
  }
  if (lane == 0) {
    H.own[2 * i] = 0;
    H.own[2 * i + 1] = 0;
    H.ovf[i] = 0;
  }
}

struct HubChunk {
  int i;              // hub (position in the batch's hub list)
  uint32_t u, o;      // unit and its module
  uint32_t ob, od, ib, j0, j1, np;
};

__device__ inline HubChunk hub_chunk(const EvalArgs &a, const HubParts &H, uint32_t c) {
  HubChunk h;
  h.i = (int)H.chub[c];
  h.u = H.hubu[h.i];
  h.o = a.mod[h.u];
  h.ob = a.optr[h.u];
  h.od = a.optr[h.u + 1] - h.ob;
  h.ib = a.iptr[h.u];
  const uint32_t deg = h.od + (a.iptr[h.u + 1] - h.ib);
  h.j0 = (c - H.A[h.i]) * kHubArcChunk;
  h.j1 = min(deg, h.j0 + kHubArcChunk);
  h.np = H.PB[h.i + 1] - H.PB[h.i];
  return h;
}

// count pass: modules of the chunk's arcs, partition sizes, own-module flows
__global__ __launch_bounds__(256) void k_hubp_count(EvalArgs a, HubParts H) {
  __shared__ uint32_t hist[kHubHist];
  __shared__ long long red[2][256 / kWave];
  const uint32_t total = H.tot[0];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const HubChunk h = hub_chunk(a, H, c);
    const bool local = h.np <= (uint32_t)kHubHist;
    if (local)
      for (uint32_t b = threadIdx.x; b < h.np; b += 256) hist[b] = 0;
    __syncthreads();
    long long so = 0, si = 0;
    uint32_t run_q = 0xffffffffu, run_n = 0;   // runs of one partition: one counter update each
    auto flush_run = [&]() {
      if (run_n) {
        if (local) atomicAdd(&hist[run_q], run_n);
        else atomicAdd(&H.bcnt[H.PB[h.i] + run_q], run_n);
      }
    };
    // four arcs per lane in flight: their ids, then their neighbours' modules
    for (uint32_t jb = h.j0 + threadIdx.x; jb < h.j1; jb += 4 * 256) {
      uint32_t kq[4], nq[4], mq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t j = jb + (uint32_t)q * 256;
        const bool out = j < h.od;
        kq[q] = out ? h.ob + j : h.ib + (j - h.od);
        nq[q] = j < h.j1 ? (out ? a.oidx[kq[q]] : a.iidx[kq[q]]) : 0u;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) mq[q] = jb + (uint32_t)q * 256 < h.j1 ? a.mod[nq[q]] : 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t j = jb + (uint32_t)q * 256;
        if (j >= h.j1) break;
        const bool out = j < h.od;
        const uint32_t m = mq[q];
        if (m == h.o) {   // own module: only its sums are needed, the arc is not partitioned
          H.mtmp[H.AB[h.i] + j] = kOwnArc;
          if (out) so += a.aout.fxw(kq[q]);
          else si += a.ain.fxw(kq[q]);
          continue;
        }
        H.mtmp[H.AB[h.i] + j] = out ? m : (m | kInBit);
        const uint32_t pq = part_of(m, h.np);
        if (pq != run_q) {
          flush_run();
          run_q = pq;
          run_n = 0;
        }
        ++run_n;
      }
    }
    flush_run();
    so = wave_sum(so);
    si = wave_sum(si);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      red[0][threadIdx.x / kWave] = so;
      red[1][threadIdx.x / kWave] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      long long to = 0, ti = 0;
      for (int q = 0; q < 256 / kWave; ++q) {
        to += red[0][q];
        ti += red[1][q];
      }
      if (to) atomicAdd((unsigned long long *)&H.own[2 * h.i], (unsigned long long)to);
      if (ti) atomicAdd((unsigned long long *)&H.own[2 * h.i + 1], (unsigned long long)ti);
    }
    if (local)
      for (uint32_t b = threadIdx.x; b < h.np; b += 256)
        if (hist[b]) atomicAdd(&H.bcnt[H.PB[h.i] + b], hist[b]);
    __syncthreads();
  }
}

// scatter pass: every arc to its partition (order inside a partition is
// irrelevant: all sums are exact integers)
__global__ __launch_bounds__(256) void k_hubp_scatter(EvalArgs a, HubParts H) {
  constexpr int PER = kHubArcChunk / 256;
  __shared__ uint32_t hcnt[kHubHist];
  __shared__ uint32_t hbase[kHubHist];
  const uint32_t total = H.tot[0];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const HubChunk h = hub_chunk(a, H, c);
    const bool local = h.np <= (uint32_t)kHubHist;
    if (local)
      for (uint32_t b = threadIdx.x; b < h.np; b += 256) hcnt[b] = 0;
    __syncthreads();
    uint32_t part[PER], rank[PER], key[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const uint32_t j = h.j0 + (uint32_t)(q * 256) + threadIdx.x;
      part[q] = 0;
      rank[q] = 0;
      key[q] = 0;
      if (j < h.j1) {
        key[q] = H.mtmp[H.AB[h.i] + j];
        if (key[q] != kOwnArc) {
          part[q] = part_of(key[q] & ~kInBit, h.np);
          rank[q] = local ? atomicAdd(&hcnt[part[q]], 1u) : atomicAdd(&H.bcur[H.PB[h.i] + part[q]], 1u);
        }
      }
    }
    __syncthreads();
    if (local)
      for (uint32_t b = threadIdx.x; b < h.np; b += 256)
        if (hcnt[b]) hbase[b] = atomicAdd(&H.bcur[H.PB[h.i] + b], hcnt[b]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const uint32_t j = h.j0 + (uint32_t)(q * 256) + threadIdx.x;
      if (j < h.j1 && key[q] != kOwnArc) {
        const bool out = j < h.od;
        const uint32_t k = out ? h.ob + j : h.ib + (j - h.od);
        const uint32_t pos = H.boff[H.PB[h.i] + part[q]] + (local ? hbase[part[q]] : 0u) + rank[q];
        H.ekey[pos] = key[q];
        H.ew[pos] = out ? a.aout.fxw(k) : a.ain.fxw(k);
      }
    }
    __syncthreads();
  }
}

// one workgroup per partition: LDS aggregation, per-wave in-place compaction
// (as in k_eval_block) and dense evaluation; best per partition
template <bool F32>
__global__ __launch_bounds__(256) void k_hubp_eval(EvalArgs a, HubParts H) {
  constexpr int NW = 256 / kWave;
  __shared__ __align__(16) uint32_t s_keys[kHubTable];
  __shared__ long long s_out[kHubTable];
  __shared__ long long s_in[kHubTable];
  const int w = threadIdx.x / kWave;
  const int wl = threadIdx.x & (kWave - 1);
  unsigned long long cand_acc = 0;   // thread 0
  const uint32_t total = H.tot[1];
  for (uint32_t p = blockIdx.x; p < total; p += gridDim.x) {
    const int i = (int)H.phub[p];
    const uint32_t u = H.hubu[i];
    const UnitLoads ul = load_unit(a, u);   // in flight while the partition aggregates
    const uint32_t e0 = H.boff[p], e1 = H.boff[p + 1];
    for (int sl = threadIdx.x; sl < kHubTable; sl += 256) {
      s_keys[sl] = kEmptyKey;
      s_out[sl] = 0;
      s_in[sl] = 0;
    }
    __syncthreads();
    // A hub's arcs concentrate on few modules (its own community above all),
    // so each lane first sums runs of equal modules in registers: one LDS
    // update per run instead of one per entry on a contended slot.
    uint32_t run_m = kEmptyKey;
    long long run_o = 0, run_i = 0;
    auto flush_run = [&]() {
      if (run_m != kEmptyKey) {
        const int sl = lds_insert_bounded(s_keys, kHubTable, run_m);
        if (sl >= 0) {
          if (run_o) atomicAdd((unsigned long long *)&s_out[sl], (unsigned long long)run_o);
          if (run_i) atomicAdd((unsigned long long *)&s_in[sl], (unsigned long long)run_i);
        }
      }
    };
    // four entries per lane in flight (an entry key never equals kEmptyKey)
    for (uint32_t eb = e0 + threadIdx.x; eb < e1; eb += 4 * 256) {
      uint32_t kq[4];
      long long wq[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t e = eb + (uint32_t)q * 256;
        kq[q] = e < e1 ? H.ekey[e] : kEmptyKey;
        wq[q] = e < e1 ? H.ew[e] : 0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (kq[q] == kEmptyKey) break;
        const uint32_t m = kq[q] & ~kInBit;
        if (m != run_m) {
          flush_run();
          run_m = m;
          run_o = 0;
          run_i = 0;
        }
        if (kq[q] & kInBit) run_i += wq[q];
        else run_o += wq[q];
      }
    }
    flush_run();
    __syncthreads();   // a failed insert leaves every slot occupied: detected below from the count
    int nw = 0;
    for (int s0 = 0; s0 < kHubTable; s0 += 256) {
      const int sl = s0 + w * kWave + wl;
      const uint32_t key = s_keys[sl];
      const long long vo = s_out[sl], vi = s_in[sl];
      const bool ne = key != kEmptyKey;
      const uint64_t bal = __ballot(ne);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (ne) {
        const int d = wave_slot(nw + (int)__popcll(bal & lanes_below()), w);
        s_keys[d] = key;
        s_out[d] = vo;
        s_in[d] = vi;
      }
      nw += (int)__popcll(bal);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    UnitCtx U;
    unit_ctx(a, ul, H.own[2 * i], H.own[2 * i + 1], U);
    Best b{0.0, kNone, -1};
    int rows = 0;
    eval_cands_bounded<true, 4, kWave, F32>(a, U, s_keys, s_out, s_in, wl, nw, ~0ull,
                                            [w](int j) { return wave_slot(j, w); }, b, &rows);
    b = group_best<kWave>(b);
    __syncthreads();   // keys are not read again: the per-wave bests reuse them
    double *r_bd = (double *)s_keys;
    uint32_t *r_bm = s_keys + 2 * NW;
    int *r_bi = (int *)(s_keys + 3 * NW);
    unsigned long long *r_nc = (unsigned long long *)(s_keys + 4 * NW);   // table slots in use
    unsigned long long *r_rows = (unsigned long long *)(s_keys + 6 * NW);   // module rows read
    if (wl == 0) {
      r_bd[w] = b.d;
      r_bm[w] = b.m;
      r_bi[w] = b.i;
      r_nc[w] = (unsigned long long)nw;
      r_rows[w] = (unsigned long long)rows;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Best bb{r_bd[0], r_bm[0], r_bi[0]};
      unsigned long long nc = r_nc[0];
      cand_acc += r_rows[0];
      for (int q = 1; q < NW; ++q) {
        nc += r_nc[q];
        cand_acc += r_rows[q];
        if (better(r_bd[q], r_bm[q], bb)) { bb.d = r_bd[q]; bb.m = r_bm[q]; bb.i = r_bi[q]; }
      }
      if (nc >= (unsigned long long)kHubTable || H.force_ovf) {
        // the table filled up (an insert may have failed): the hub goes to the fallback
        H.ovf[i] = 1u;
        H.bd[p] = 0.0;
        H.bm[p] = kNone;
        atomicAdd(&a.counters[cHubOvf], 1u);
      } else {
        H.bd[p] = bb.d;
        H.bm[p] = bb.m;
        H.bfo[p] = bb.i >= 0 ? s_out[bb.i] : 0;
        H.bfi[p] = bb.i >= 0 ? s_in[bb.i] : 0;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && cand_acc) add_cands(a, cand_acc);
}

// one wavefront per hub: best over its partitions (fixed order), empty
// option, emit; hubs flagged for overflow are left to the fallback
__global__ __launch_bounds__(256) void k_hubp_finish(EvalArgs a, HubParts H) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t wstride = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = w0; i < a.list_n; i += wstride) {
    if (H.ovf[i]) continue;
    const uint32_t u = a.bucket[a.list[i]];
    Best b{0.0, kNone, -1};
    for (uint32_t p = H.PB[i] + lane; p < H.PB[i + 1]; p += kWave)
      if (better(H.bd[p], H.bm[p], b)) { b.d = H.bd[p]; b.m = H.bm[p]; b.i = (int)p; }
    b = group_best<kWave>(b);
    if (lane == 0) {
      const long long fo_o = H.own[2 * i], fi_o = H.own[2 * i + 1];
      UnitCtx U;
      unit_ctx(a, u, fo_o, fi_o, U);
      if (U.offer_empty) {
        const double dl = delta_empty(U);
        if (better(dl, a.empty_key, b) && dl < 0.0) { b.d = dl; b.m = a.empty_key; b.i = -1; }
      }
      const long long fo_n = b.i >= 0 ? H.bfo[b.i] : 0, fi_n = b.i >= 0 ? H.bfi[b.i] : 0;
      emit(a, u, b, fo_o, fi_o, fo_n, fi_n);
      add_visits(a, (unsigned long long)((a.optr[u + 1] - a.optr[u]) + (a.iptr[u + 1] - a.iptr[u])));
      add_units(a, 1u);
    }
  }
}

// ------------------------------------------------------------------ commit
// Final target of every mover: a neighbour module, an empty module by rank
// among this batch's empty requests (ascending unit id), or invalid when no
// empty module is left (Module.cpp:1138-1152).
// Empty-module requests of a batch: one bit per unit (set by the evaluation,
// cleared once the batch is resolved).  A request's rank among the batch's
// requests, in unit order, is the requests of the words before its own plus
// those of the lower bits of its word: want_pre holds the exclusive scan of
// the words' bit counts, 32 times shorter than a scan over the units.
__device__ inline uint32_t want_rank(const uint32_t *__restrict__ want, const uint32_t *__restrict__ want_pre,
                                     uint32_t u) {
  return want_pre[u >> 5] + (uint32_t)__popc(want[u >> 5] & ((1u << (u & 31u)) - 1u));
}

__global__ void k_popc_words(const uint32_t *__restrict__ w, int64_t nw, uint32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) out[i] = (uint32_t)__popc(w[i]);
}

__global__ void k_resolve(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                          const uint32_t *__restrict__ target, const uint32_t *__restrict__ want,
                          const uint32_t *__restrict__ want_pre,
                          const uint32_t *__restrict__ empty_list, uint32_t empty_cursor, uint32_t n_empty,
                          uint32_t empty_key, uint32_t *__restrict__ newmod, uint32_t *__restrict__ stamp,
                          uint32_t cur) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    uint32_t t = target[u];
    if (t == empty_key) {
      const uint32_t r = empty_cursor + want_rank(want, want_pre, u);
      t = r < n_empty ? empty_list[r] : kInvalid;
    }
    newmod[u] = t;
    if (t != kInvalid) stamp[u] = cur;
  }
}

// The same for a batch: an empty-module request is served from the
// requester's instance, by its rank among that instance's requests: its rank
// among the level's requests (units grouped by instance) less that of the
// instance's first unit; *total counts all requests.
__global__ void k_resolve_batch(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                                const uint32_t *__restrict__ target, const uint32_t *__restrict__ want,
                                const uint32_t *__restrict__ want_pre,
                                const uint32_t *__restrict__ total, int32_t n_units,
                                const uint32_t *__restrict__ uinst, const uint32_t *__restrict__ uoff,
                                const uint32_t *__restrict__ empty_list, const uint32_t *__restrict__ e_beg,
                                const uint32_t *__restrict__ e_cnt, const uint32_t *__restrict__ e_cur,
                                uint32_t empty_key, uint32_t *__restrict__ newmod, uint32_t *__restrict__ stamp,
                                uint32_t cur) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    uint32_t t = target[u];
    if (t == empty_key) {
      const uint32_t I = uinst[u];
      const uint32_t first = uoff[I];
      const uint32_t before = (int32_t)first < n_units ? want_rank(want, want_pre, first) : *total;
      const uint32_t r = e_cur[I] + (want_rank(want, want_pre, u) - before);
      t = r < e_cnt[I] ? empty_list[e_beg[I] + r] : kInvalid;
    }
    newmod[u] = t;
    if (t != kInvalid) stamp[u] = cur;
  }
}

// after k_resolve_batch: each instance's taken empty modules
__global__ void k_bempty_advance(int32_t nb, const uint32_t *__restrict__ uoff, const uint32_t *__restrict__ want,
                                 const uint32_t *__restrict__ want_pre, const uint32_t *__restrict__ total,
                                 int32_t n_units, const uint32_t *__restrict__ e_cnt, uint32_t *__restrict__ e_cur) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  const uint32_t f = uoff[I], l = uoff[I + 1];
  const uint32_t a = (int32_t)f < n_units ? want_rank(want, want_pre, f) : *total;
  const uint32_t b = (int32_t)l < n_units ? want_rank(want, want_pre, l) : *total;
  const uint32_t req = b - a, left = e_cnt[I] - e_cur[I];
  e_cur[I] += req < left ? req : left;
}

// every instance's segment of the (ascending) empty-module list: the ids in
// [off[I], off[I + 1])
__global__ void k_bempty_seg(int32_t nb, const uint32_t *__restrict__ off, const uint32_t *__restrict__ empty_list,
                             uint32_t n_empty, uint32_t *__restrict__ e_beg, uint32_t *__restrict__ e_cnt,
                             uint32_t *__restrict__ e_cur) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  auto lower = [&](uint32_t v) {
    uint32_t lo = 0, hi = n_empty;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (empty_list[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const uint32_t b = lower(off[I]), e = lower(off[I + 1]);
  e_beg[I] = b;
  e_cnt[I] = e - b;
  e_cur[I] = 0;
}

// the heads idx[b..e) of one unit's arcs become active (G lanes, kRowAhead
// heads per lane in flight)
template <int G>
__device__ inline void mark_heads(const uint32_t *__restrict__ idx, uint32_t b, uint32_t e, int lane,
                                  uint8_t *act_next) {
  for (uint32_t k0 = b + lane; k0 < e; k0 += kRowAhead * G) {
    uint32_t v[kRowAhead];
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < e ? idx[k0 + j * G] : 0u;
#pragma unroll
    for (int j = 0; j < kRowAhead; ++j)
      if (k0 + j * G < e) act_next[v[j]] = 1;
  }
}

// G lanes per mover: module statistics, exit flows and activity marks.
// X changes of a move of u from o to n with every neighbour fixed:
//   X_o += -(W_u - A_o) + B_o,  X_n += (W_u - A_n) - B_n;
// arcs u->v whose head moves in the same batch are corrected from the tail:
//   X_o += w ([o != o_v] - [o != n_v]),  X_n += w ([n != n_v] - [n != o_v]).
// The arcs of movers with more than big_deg arcs are left to k_commit_big,
// which spreads them over many workgroups (all updates are exact integer
// atomics, so the split does not change the result).
// Activity marks (mark != 0): every mover's out- and in-neighbours become
// active in the next sweep (Module.cpp:1661-1668), movers rejected for want
// of an empty module included.  A batch with at least pull_min movers only
// flags its movers in wmark and sets *pull_flag; k_pull_marks then marks
// their neighbours at the end of the sweep from the receiving side, where a
// unit stops at its first flagged neighbour.
template <int G>
__global__ __launch_bounds__(256) void k_commit(const uint32_t *__restrict__ movers,
                                                const uint32_t *__restrict__ counters, const uint32_t *__restrict__ optr,
                                                const uint32_t *__restrict__ oidx, const uint32_t *__restrict__ iptr,
                                                const uint32_t *__restrict__ iidx, ArcW aout,
                                                const uint32_t *__restrict__ mod, const uint32_t *__restrict__ newmod,
                                                const uint32_t *__restrict__ stamp, uint32_t cur,
                                                const long long *__restrict__ agg, const Stat4 *__restrict__ us,
                                                const int32_t *__restrict__ nmem, Stat4 *ms, int32_t *mNodes,
                                                int32_t *mUnits, uint8_t *act_next, int mark, uint32_t *stripes,
                                                uint32_t big_deg, uint32_t *big_list, uint32_t *big_cnt,
                                                uint32_t pull_min, uint8_t *wmark, uint32_t *pull_flag,
                                                unsigned long long *smax) {
  constexpr int GPB = 256 / G;   // movers per workgroup and round
  // Each round takes GPB movers; their module-row updates (old and new module
  // of each) are merged by module in LDS before the atomics, so that movers
  // of one round joining the same module update its row once.
  __shared__ uint32_t s_mod[2 * GPB];
  __shared__ long long s_dv[2 * GPB][4];   // flow, teleport, dangling flow, exit flow
  __shared__ int32_t s_dn[2 * GPB];        // member nodes (units: +-1)
  const uint32_t nm = counters[cMovers];
  const bool pull = mark && nm >= pull_min;
  if (pull && blockIdx.x == 0 && threadIdx.x == 0) *pull_flag = 1u;
  const int lane = threadIdx.x % G;
  const int gl = threadIdx.x / G;
  uint32_t moved_here = 0;
  unsigned long long grown[2] = {0, 0};   // largest flow and teleport weight a module row reached here
  for (int64_t r0 = (int64_t)blockIdx.x * GPB; r0 < nm; r0 += (int64_t)gridDim.x * GPB) {
    const int64_t i = r0 + gl;
    bool live = i < nm;   // this group's mover commits a row update this round
    if (live) {
      const uint32_t u = movers[i];
      const uint32_t ob = optr[u], oe = optr[u + 1];
      const bool big = (oe - ob) + (iptr[u + 1] - iptr[u]) > big_deg;
      if (pull && lane == 0) wmark[u] = 1;
      if (big) {
        if (lane == 0) big_list[atomicAdd(big_cnt, 1u)] = u;
      } else if (mark && !pull) {
        mark_heads<G>(oidx, ob, oe, lane, act_next);
        mark_heads<G>(iidx, iptr[u], iptr[u + 1], lane, act_next);
      }
      live = stamp[u] == cur;
      if (live) {
        const uint32_t o = mod[u], n = newmod[u];
        long long dxo = 0, dxn = 0;
        // kRowAhead arcs per lane in flight: heads, then their stamps
        const uint32_t oend = big ? ob : oe;
        for (uint32_t k0 = ob + lane; k0 < oend; k0 += kRowAhead * G) {
          uint32_t v[kRowAhead];
          bool co[kRowAhead];
#pragma unroll
          for (int j = 0; j < kRowAhead; ++j) v[j] = k0 + j * G < oend ? oidx[k0 + j * G] : 0u;
#pragma unroll
          for (int j = 0; j < kRowAhead; ++j) co[j] = k0 + j * G < oend && stamp[v[j]] == cur;
#pragma unroll
          for (int j = 0; j < kRowAhead; ++j) {
            if (!co[j]) continue;
            const long long w = aout.fxw(k0 + j * G);
            const uint32_t ov = mod[v[j]], nv = newmod[v[j]];
            dxo += w * ((long long)(o != ov) - (long long)(o != nv));
            dxn += w * ((long long)(n != nv) - (long long)(n != ov));
          }
        }
        dxo = group_sum_ll<G>(dxo);
        dxn = group_sum_ll<G>(dxn);
        if (lane == 0) {
          const long long *g = agg + 4 * (int64_t)u;
          const Stat4 su = us[u];
          const long long W = su.x;
          dxo += -(W - g[0]) + g[1];
          dxn += (W - g[2]) - g[3];
          const int32_t k = nmem[u];
          s_mod[2 * gl] = o;
          s_dv[2 * gl][0] = -su.p;
          s_dv[2 * gl][1] = -su.t;
          s_dv[2 * gl][2] = -su.d;
          s_dv[2 * gl][3] = dxo;
          s_dn[2 * gl] = -k;
          s_mod[2 * gl + 1] = n;
          s_dv[2 * gl + 1][0] = su.p;
          s_dv[2 * gl + 1][1] = su.t;
          s_dv[2 * gl + 1][2] = su.d;
          s_dv[2 * gl + 1][3] = dxn;
          s_dn[2 * gl + 1] = k;
          moved_here += (uint32_t)k;
        }
      }
    }
    if (!live && lane == 0) {
      s_mod[2 * gl] = kNone;
      s_mod[2 * gl + 1] = kNone;
    }
    __syncthreads();
    // entry t: the first entry of its module adds every entry of that module
    const int t = threadIdx.x;
    if (t < 2 * GPB && s_mod[t] != kNone) {
      const uint32_t m = s_mod[t];
      bool first = true;
      for (int j = 0; j < t; ++j) first = first && s_mod[j] != m;
      if (first) {
        long long dv[4] = {s_dv[t][0], s_dv[t][1], s_dv[t][2], s_dv[t][3]};
        int32_t dn = s_dn[t], du = (t & 1) ? 1 : -1;
        for (int j = t + 1; j < 2 * GPB; ++j) {
          if (s_mod[j] != m) continue;
#pragma unroll
          for (int f = 0; f < 4; ++f) dv[f] += s_dv[j][f];
          dn += s_dn[j];
          du += (j & 1) ? 1 : -1;
        }
        unsigned long long *row = (unsigned long long *)&ms[m];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          if (!dv[f]) continue;
          const unsigned long long old = atomicAdd(row + f, (unsigned long long)dv[f]);
          if (f < 2 && dv[f] > 0) grown[f] = max(grown[f], old + (unsigned long long)dv[f]);
        }
        if (dn) atomicAdd(&mNodes[m], dn);
        if (du) atomicAdd(&mUnits[m], du);
      }
    }
    __syncthreads();
  }
  moved_here = wave_sum(moved_here);
  if ((threadIdx.x & (kWave - 1)) == 0 && moved_here) atomicAdd(my_stripe(stripes) + cMoved, moved_here);
  // the candidate bound's module maxima (one atomic per wavefront at most)
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    grown[0] = max(grown[0], (unsigned long long)__shfl_xor(grown[0], off, kWave));
    grown[1] = max(grown[1], (unsigned long long)__shfl_xor(grown[1], off, kWave));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    // Stat4 holds p then t; smax holds T then P
    if (grown[1] > smax[0]) atomicMax(&smax[0], grown[1]);
    if (grown[0] > smax[1]) atomicMax(&smax[1], grown[0]);
  }
}

// Profiling only: arcs of a batch's movers, split at big_deg: out-arcs and
// all arcs of the movers at or below it (k_commit's), then of those above
// it (k_commit_big's).
__global__ __launch_bounds__(256) void k_mover_arcs(const uint32_t *__restrict__ movers,
                                                    const uint32_t *__restrict__ counters,
                                                    const uint32_t *__restrict__ optr,
                                                    const uint32_t *__restrict__ iptr, uint32_t big_deg,
                                                    unsigned long long *__restrict__ out) {
  const uint32_t nm = counters[cMovers];
  unsigned long long a[4] = {0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    const uint32_t od = optr[u + 1] - optr[u];
    const uint32_t d = od + (iptr[u + 1] - iptr[u]);
    const int hi = d > big_deg ? 2 : 0;
    a[hi] += od;
    a[hi + 1] += d;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const unsigned long long v = wave_sum(a[k]);
    if ((threadIdx.x & (kWave - 1)) == 0 && v) atomicAdd(&out[k], v);
  }
}

// Arc chunks of the deferred (high-degree) movers: prefix of chunk counts.
constexpr int kCommitChunk = 4096;
// smallest batch (units) whose marks may be pulled
constexpr uint32_t kPullMinUnits = 65536;

__global__ __launch_bounds__(256) void k_commit_plan(const uint32_t *__restrict__ big_list,
                                                     const uint32_t *__restrict__ big_cnt,
                                                     const uint32_t *__restrict__ optr,
                                                     const uint32_t *__restrict__ iptr, uint32_t *__restrict__ A) {
  __shared__ uint32_t lds[256 / kWave];
  const int cnt = (int)*big_cnt;
  uint32_t carry = 0;
  if (threadIdx.x == 0) A[0] = 0;
  for (int base = 0; base < cnt; base += 256) {
    const int i = base + threadIdx.x;
    uint32_t ac = 0;
    if (i < cnt) {
      const uint32_t u = big_list[i];
      const uint32_t deg = (optr[u + 1] - optr[u]) + (iptr[u + 1] - iptr[u]);
      ac = (deg + kCommitChunk - 1) / kCommitChunk;
    }
    uint32_t t;
    const uint32_t inc = block_incl_scan(ac, lds, &t);
    if (i < cnt) A[i + 1] = carry + inc;
    carry += t;
  }
}

// one workgroup per arc chunk of a deferred mover: activity marks (unless
// the batch leaves them to k_pull_marks) and the co-mover corrections of its
// out-arcs
__global__ __launch_bounds__(256) void k_commit_big(const uint32_t *__restrict__ big_list,
                                                    const uint32_t *__restrict__ big_cnt,
                                                    const uint32_t *__restrict__ A, const uint32_t *__restrict__ optr,
                                                    const uint32_t *__restrict__ oidx, const uint32_t *__restrict__ iptr,
                                                    const uint32_t *__restrict__ iidx, ArcW aout,
                                                    const uint32_t *__restrict__ mod,
                                                    const uint32_t *__restrict__ newmod,
                                                    const uint32_t *__restrict__ stamp, uint32_t cur, Stat4 *ms,
                                                    uint8_t *act_next, int mark_in,
                                                    const uint32_t *__restrict__ counters, uint32_t pull_min) {
  __shared__ long long lds[2][256 / kWave];
  const int cnt = (int)*big_cnt;
  if (cnt == 0) return;
  const bool mark = mark_in && counters[cMovers] < pull_min;
  const uint32_t total = A[cnt];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const int i = chunk_owner(A, cnt, c);
    const uint32_t u = big_list[i];
    const uint32_t ob = optr[u], od = optr[u + 1] - ob;
    const uint32_t ib = iptr[u], deg = od + (iptr[u + 1] - ib);
    const uint32_t j0 = (c - A[i]) * kCommitChunk;
    const uint32_t j1 = min(deg, j0 + kCommitChunk);
    const bool moves = stamp[u] == cur;
    const uint32_t o = mod[u], n = newmod[u];
    long long dxo = 0, dxn = 0;
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += 256) {
      if (j < od) {
        const uint32_t k = ob + j;
        const uint32_t v = oidx[k];
        if (mark) act_next[v] = 1;
        if (moves && stamp[v] == cur) {
          const long long w = aout.fxw(k);
          const uint32_t ov = mod[v], nv = newmod[v];
          dxo += w * ((long long)(o != ov) - (long long)(o != nv));
          dxn += w * ((long long)(n != nv) - (long long)(n != ov));
        }
      } else if (mark) {
        act_next[iidx[ib + (j - od)]] = 1;
      }
    }
    dxo = wave_sum(dxo);
    dxn = wave_sum(dxn);
    if ((threadIdx.x & (kWave - 1)) == 0) {
      lds[0][threadIdx.x / kWave] = dxo;
      lds[1][threadIdx.x / kWave] = dxn;
    }
    __syncthreads();
    if (threadIdx.x == 0 && moves) {
      long long so = 0, sn = 0;
      for (int q = 0; q < 256 / kWave; ++q) {
        so += lds[0][q];
        sn += lds[1][q];
      }
      if (so) atomicAdd((unsigned long long *)&ms[o].x, (unsigned long long)so);
      if (sn) atomicAdd((unsigned long long *)&ms[n].x, (unsigned long long)sn);
    }
    __syncthreads();
  }
}

template <int G>
__device__ inline int group_or(int v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off, G);
  return v;
}

// whether flag[idx[k]] is set for some k in [b, e): G lanes, one arc per
// lane in the first round and twice as many in each next one (up to 8), up to
// the first round that finds one
template <int G>
__device__ inline bool any_flagged(const uint32_t *__restrict__ idx, uint32_t b, uint32_t e,
                                   const uint8_t *__restrict__ flag, int lane) {
  int per = 1;
  for (uint32_t k0 = b; k0 < e; k0 += (uint32_t)(per * G), per = per < 8 ? 2 * per : 8) {
    int h = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t k = k0 + (uint32_t)(j * G + lane);
      if (j < per && k < e) h |= flag[idx[k]];
    }
    if (group_or<G>(h)) return true;
  }
  return false;
}

// End of a sweep in which some batch left its activity marks to the
// receiving side (*pull_flag set by k_commit): a unit not yet active for the
// next sweep becomes active when an out- or in-neighbour is flagged in
// wmark.  With most units moving, a unit usually stops at its first
// neighbour, where pushing marks writes one byte per arc of every mover.
template <int G>
__global__ __launch_bounds__(256) void k_pull_marks(const uint32_t *__restrict__ pull_flag, int32_t n,
                                                    const uint32_t *__restrict__ optr,
                                                    const uint32_t *__restrict__ oidx,
                                                    const uint32_t *__restrict__ iptr,
                                                    const uint32_t *__restrict__ iidx,
                                                    const uint8_t *__restrict__ wmark, uint8_t *__restrict__ act_next) {
  if (*pull_flag == 0u) return;
  const int lane = threadIdx.x % G;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x / G;
  for (int64_t u = g0; u < n; u += gstride) {
    if (act_next[u]) continue;
    const bool hit = any_flagged<G>(oidx, optr[u], optr[u + 1], wmark, lane) ||
                     any_flagged<G>(iidx, iptr[u], iptr[u + 1], wmark, lane);
    if (hit && lane == 0) act_next[u] = 1;
  }
}

// after a batch's commits, before k_apply: q of the old and the new module
// of every mover (a module met twice gets the same value twice)
__global__ void k_qv_touched(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                             const uint32_t *__restrict__ stamp, uint32_t cur, const uint32_t *__restrict__ mod,
                             const uint32_t *__restrict__ newmod, const Stat4 *__restrict__ ms,
                             const int32_t *__restrict__ mNodes, double *__restrict__ qv) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    if (stamp[u] != cur) continue;
    const uint32_t o = mod[u], n = newmod[u];
    const Stat4 so = ms[o], sn = ms[n];
    qv[o] = mNodes[o] > 0 ? module_exit(so.p, so.t, so.d, so.x) : 0.0;
    qv[n] = mNodes[n] > 0 ? module_exit(sn.p, sn.t, sn.d, sn.x) : 0.0;
  }
}

// Every writer of bit m sets it to the same value (from m's final count), so
// a bit that already holds it needs no atomic: movers joining one popular
// module then do not all update the same word.
__device__ inline void set_one(uint32_t *one, const int32_t *mUnits, uint32_t m) {
  const uint32_t bit = 1u << (m & 31u);
  const bool want = mUnits[m] == 1;
  if (((one[m >> 5] & bit) != 0) == want) return;
  if (want) atomicOr(&one[m >> 5], bit);
  else atomicAnd(&one[m >> 5], ~bit);
}

// New labels of the batch's movers, and the one-unit bits of the modules the
// commits changed (every mover's old and new module; the commits are done,
// so each bit is set from its module's final count)
__global__ void k_apply(const uint32_t *__restrict__ movers, const uint32_t *__restrict__ counters,
                        const uint32_t *__restrict__ newmod, const uint32_t *__restrict__ stamp, uint32_t cur,
                        uint32_t *__restrict__ mod, const int32_t *__restrict__ mUnits, uint32_t *one) {
  const uint32_t nm = counters[cMovers];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t u = movers[i];
    if (stamp[u] != cur) continue;
    const uint32_t o = mod[u], n = newmod[u];
    mod[u] = n;
    set_one(one, mUnits, o);
    set_one(one, mUnits, n);
  }
}

// one-unit bits of modules [0, n), and with stale counts (node level) their
// more-than-one-node bits.  One module per lane (coalesced reads); a
// wavefront's 64 bits are its ballot, written as two words by lanes 0 and 1.
// Launched with 256-lane blocks, so every wavefront starts at a multiple of
// 64 modules.
__global__ __launch_bounds__(256) void k_one_bits(const int32_t *__restrict__ mUnits,
                                                  const int32_t *__restrict__ mStale, int32_t n,
                                                  uint32_t *__restrict__ one, uint32_t *__restrict__ multi) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w0 = (m - lane) / 32;   // first word of this wavefront's modules
  if (w0 * 32 >= n) return;             // the whole wavefront lies past the end
  const uint64_t b1 = __ballot(m < n && mUnits[m] == 1);
  const uint64_t b2 = mStale ? __ballot(m < n && mStale[m] > 1) : 0ull;
  const int64_t words = ((int64_t)n + 31) / 32;
  if (lane < 2 && w0 + lane < words) {
    one[w0 + lane] = (uint32_t)(b1 >> (32 * lane));
    if (mStale) multi[w0 + lane] = (uint32_t)(b2 >> (32 * lane));
  }
}

// sort key of every bucket position: (bucket, batch) or (bucket, inactive)
__global__ void k_batch_keys(const uint32_t *__restrict__ bucket, int32_t n, int4 off_lo, int4 off_hi,
                             const uint8_t *__restrict__ act, int all_active, uint64_t seed,
                             const uint32_t *__restrict__ uinst, const uint64_t *__restrict__ iseed, int batches,
                             uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int offs[8] = {off_lo.x, off_lo.y, off_lo.z, off_lo.w, off_hi.x, off_hi.y, off_hi.z, off_hi.w};
  int b = 0;
  while (b + 1 < kNumBuckets && i >= offs[b + 1]) ++b;
  const uint32_t u = bucket[i];
  const uint64_t sd = uinst ? seed ^ iseed[uinst[u]] : seed;   // a batch: every instance its own draw
  const uint32_t k = (all_active || act[u]) ? batch_of(sd, batches, u) : (uint32_t)batches;
  keys[i] = (uint32_t)b * (uint32_t)(batches + 1) + k;
  vals[i] = b + 1 < kNumBuckets ? u : (uint32_t)i;   // hubs keep their bucket position (their table region)
}

__global__ void k_hist256(const uint32_t *__restrict__ key, int64_t n, uint32_t *__restrict__ cnt) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[key[i] & 255u], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

// Appends the movers held by the counter copies to the dense mover list,
// after the movers already there (counters[cMovers]).
__global__ __launch_bounds__(256) void k_gather_movers(const uint32_t *__restrict__ counters,
                                                       const uint32_t *__restrict__ stripes,
                                                       const uint32_t *__restrict__ mv_st, uint32_t cap,
                                                       uint32_t *__restrict__ movers) {
  __shared__ uint32_t pre[kStripes + 1];
  if (threadIdx.x < kStripes) pre[threadIdx.x + 1] = min(stripes[threadIdx.x * kStripeWords + cMovers], cap);
  if (threadIdx.x == 0) pre[0] = 0;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = 0; j < kStripes; ++j) pre[j + 1] += pre[j];
  __syncthreads();
  const uint32_t base = counters[cMovers], total = pre[kStripes];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int lo = 0, hi = kStripes;   // pre[lo] <= i < pre[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid] <= i) lo = mid;
      else hi = mid;
    }
    movers[base + i] = mv_st[(size_t)lo * cap + (i - pre[lo])];
  }
}

// Adds the counter copies to `counters` (one workgroup of kStripes lanes,
// after k_gather_movers) and clears them.
__global__ __launch_bounds__(kStripes) void k_fold_stripes(uint32_t *__restrict__ counters,
                                                           uint32_t *__restrict__ stripes, uint32_t cap) {
  constexpr int NW = kStripes / kWave;
  __shared__ unsigned long long part[6][NW];
  uint32_t *st = stripes + threadIdx.x * kStripeWords;
  unsigned long long v[6] = {min(st[cMovers], cap), st[cWant], st[cMoved],
                             *(const unsigned long long *)(st + cVisitsLo),
                             *(const unsigned long long *)(st + cCandsLo), st[cUnits]};
  st[cMovers] = 0;
  st[cWant] = 0;
  st[cMoved] = 0;
  *(unsigned long long *)(st + cVisitsLo) = 0;
  *(unsigned long long *)(st + cCandsLo) = 0;
  st[cUnits] = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[k] = wave_sum(v[k]);
    if ((threadIdx.x & (kWave - 1)) == 0) part[k][threadIdx.x / kWave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 6; ++k)
      for (int w = 0; w < NW; ++w) t[k] += part[k][w];
    counters[cMovers] += (uint32_t)t[0];
    counters[cWant] += (uint32_t)t[1];
    counters[cMoved] += (uint32_t)t[2];
    *(unsigned long long *)(counters + cVisitsLo) += t[3];
    *(unsigned long long *)(counters + cCandsLo) += t[4];
    counters[cUnits] += (uint32_t)t[5];
  }
}

__global__ void k_reset_counters(uint32_t *counters) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0;
  if (threadIdx.x == 2) counters[cBig] = 0;
  if (threadIdx.x == 3) counters[cHubOvf] = 0;
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

// The evaluation kernel of degree bucket bk (below the hubs) over a.list.
void launch_bucket(int bk, const EvalArgs &a, int32_t cnt, bool f32, hipStream_t s) {
  switch (bk) {
    case 0:
      if (f32) k_eval_group<8, 32, true><<<(unsigned)((cnt + 31) / 32), 256, 0, s>>>(a);
      else k_eval_group<8, 32, false><<<(unsigned)((cnt + 31) / 32), 256, 0, s>>>(a);
      break;
    case 1:
      if (f32) k_eval_group<16, 64, true><<<(unsigned)((cnt + 15) / 16), 256, 0, s>>>(a);
      else k_eval_group<16, 64, false><<<(unsigned)((cnt + 15) / 16), 256, 0, s>>>(a);
      break;
    case 2:
      if (f32) k_eval_group<32, 256, true><<<(unsigned)((cnt + 7) / 8), 256, 0, s>>>(a);
      else k_eval_group<32, 256, false><<<(unsigned)((cnt + 7) / 8), 256, 0, s>>>(a);
      break;
    case 3:
      if (f32) k_eval_group<64, 512, true><<<(unsigned)((cnt + 3) / 4), 256, 0, s>>>(a);
      else k_eval_group<64, 512, false><<<(unsigned)((cnt + 3) / 4), 256, 0, s>>>(a);
      break;
    case 4:
      if (f32) k_eval_group<64, 1024, true><<<(unsigned)((cnt + 3) / 4), 256, 0, s>>>(a);
      else k_eval_group<64, 1024, false><<<(unsigned)((cnt + 3) / 4), 256, 0, s>>>(a);
      break;
    case 5:
      if (f32) k_eval_block<2048, true><<<(unsigned)(cnt < 8192 ? cnt : 8192), 256, 0, s>>>(a);
      else k_eval_block<2048, false><<<(unsigned)(cnt < 8192 ? cnt : 8192), 256, 0, s>>>(a);
      break;
    case 6:
      if (f32) k_eval_block<4096, true><<<(unsigned)(cnt < 8192 ? cnt : 8192), 256, 0, s>>>(a);
      else k_eval_block<4096, false><<<(unsigned)(cnt < 8192 ? cnt : 8192), 256, 0, s>>>(a);
      break;
    default:
      break;
  }
}

// Arguments of the evaluation kernels for one sweep of level L.
// Largest teleport weight and flow of any module row (fixed point), into
// smax[0] and smax[1] (zeroed by the caller).
__global__ __launch_bounds__(256) void k_stat_max(const Stat4 *__restrict__ ms, int64_t n,
                                                  unsigned long long *smax) {
  __shared__ unsigned long long red[2][256 / kWave];
  unsigned long long tm = 0, pm = 0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 4 * stride) {
    long long t[4], p[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + q * stride;
      t[q] = i < n ? ms[i].t : 0;
      p[q] = i < n ? ms[i].p : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      tm = max(tm, (unsigned long long)t[q]);
      pm = max(pm, (unsigned long long)p[q]);
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    tm = max(tm, (unsigned long long)__shfl_xor(tm, off, kWave));
    pm = max(pm, (unsigned long long)__shfl_xor(pm, off, kWave));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[0][threadIdx.x / kWave] = tm;
    red[1][threadIdx.x / kWave] = pm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 256 / kWave; ++w) {
      tm = max(tm, red[0][w]);
      pm = max(pm, red[1][w]);
    }
    if (tm) atomicMax(&smax[0], tm);
    if (pm) atomicMax(&smax[1], pm);
  }
}

EvalArgs sweep_args(Ctx &c, Level &L, bool prioritized, double vthresh, uint64_t sweep_seed, bool batch) {
  EvalArgs a{};
  a.bucket = L.bucket.p;
  a.hub_base = L.bucket_off[kNumBuckets - 1];
  a.optr = L.out_ptr.p;
  a.oidx = L.out_idx.p;
  a.iptr = L.in_ptr.p;
  a.iidx = L.in_idx.p;
  a.aout = arcw_out(c, L);
  a.ain = arcw_in(c, L);
  a.mod = L.mod.p;
  a.us = L.us.p;
  a.ms = c.ms.p;
  {   // the one-unit bits from the module counts as the sweep starts, and
      // at the node level the more-than-one-node bits (the stale counts do
      // not change during a sweep)
    const int64_t words = ((int64_t)c.N + 31) / 32;
    const bool node = c.level == 0;
    c.mOne.alloc(words);
    if (node) c.mMulti.alloc(words);
    k_one_bits<<<grid_for(c.N, 256), 256, 0, c.stream>>>(c.mUnits.p, node ? c.mStale.p : nullptr, c.N, c.mOne.p,
                                                         node ? c.mMulti.p : nullptr);
    RMX_LAUNCH_CHECK();
    a.mOne = c.mOne.p;
    a.mMulti = node ? c.mMulti.p : nullptr;
  }
  {   // upper bounds of the module rows' T and P for the candidate bound;
      // the commits raise them as modules grow
    c.smax.alloc(2);
    fill_i64((long long *)c.smax.p, 0, 2, c.stream);
    if (c.ms.n > 0) {
      k_stat_max<<<(unsigned)std::min<int64_t>(512, grid_for(c.ms.n, 256)), 256, 0, c.stream>>>(c.ms.p, c.ms.n,
                                                                                              c.smax.p);
      RMX_LAUNCH_CHECK();
    }
    a.smax = c.smax.p;
  }
  a.act = c.act.p;
  a.all_active = prioritized ? 0 : 1;
  a.seed = sweep_seed;
  a.batches = sweep_batches(c, L);
  a.node_level = c.level == 0 ? 1 : 0;
  a.empty_key = (uint32_t)c.N + 1u;
  a.scal = c.scal.p;
  a.S_inst = c.sub_mode ? c.inst_S.p : (batch ? c.b_S.p : nullptr);
  a.inst = c.sub_mode ? c.sub_par.p : (batch ? L.uinst.p : nullptr);
  a.e_cnt = batch ? c.b_ecnt.p : nullptr;
  a.e_cur = batch ? c.b_ecur.p : nullptr;
  a.vthresh = prioritized ? vthresh : 0.0;
  a.target = c.target.p;
  a.agg = c.agg_buf.p;
  a.movers_st = c.movers_st.p;
  a.mv_cap = c.mv_cap;
  a.counters = c.counters.p;
  a.stripes = c.counters.p + kStripeBase;
  a.want = c.want_empty.p;
  a.hkeys = c.hub_keys.p;
  a.hout = c.hub_out.p;
  a.hin = c.hub_in.p;
  a.hub_off = L.hub_off.p;
  return a;
}

// degree classes of the autotuner (upper bounds; above the last: hubs)
constexpr int kNumClasses = 9;
constexpr int kClassMax[kNumClasses] = {8, 16, 32, 64, 128, 256, 512, 1023, 2048};

__global__ void k_class_key(const uint32_t *__restrict__ optr, const uint32_t *__restrict__ iptr, int32_t n,
                            uint32_t *__restrict__ key, uint32_t *__restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t deg = (optr[i + 1] - optr[i]) + (iptr[i + 1] - iptr[i]);
  uint32_t k = 0;
  while (k < (uint32_t)kNumClasses && deg > (uint32_t)kClassMax[k]) ++k;
  key[i] = k;
  vals[i] = (uint32_t)i;
}

int64_t want_words(int64_t n) { return n / 32 + 1; }

// want_pre (in c.rank_tmp) = exclusive scan of the request words' bit counts;
// *total (device, may be null) = the number of requests
void want_prefix(Ctx &c, int64_t n, uint32_t *total, hipStream_t s) {
  const int64_t nw = want_words(n);
  k_popc_words<<<grid_for(nw, 256), 256, 0, s>>>(c.want_empty.p, nw, c.rank_tmp.p);
  RMX_LAUNCH_CHECK();
  exclusive_scan_u32(c.rank_tmp.p, c.rank_tmp.p, nw, total, c.scratch, s);
}

unsigned stride_grid(int64_t items, int per_block) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

// Folds the counter copies into the counters (movers into the dense list).
void fold_counters(Ctx &c, hipStream_t s) {
  k_gather_movers<<<512, 256, 0, s>>>(c.counters.p, c.counters.p + kStripeBase, c.movers_st.p, c.mv_cap,
                                      c.movers.p);
  RMX_LAUNCH_CHECK();
  k_fold_stripes<<<1, kStripes, 0, s>>>(c.counters.p, c.counters.p + kStripeBase, c.mv_cap);
  RMX_LAUNCH_CHECK();
}

}  // namespace

// Algorithmic bytes of one evaluation launch: per arc visit the neighbour id,
// its weight and its module (4 + w + 4 bytes); per unit its statistics, module
// and CSR bounds (48 bytes); per candidate module its four statistics (32).
static void profile_eval(Ctx &c, int bucket, hipStream_t s) {
  // entries carry the kernels' own names (both levels together); the hub
  // bucket is a sequence of kernels and is marked as such
  // the kernels' own names, with the precision template argument
  static const char *names[2][kNumBuckets] = {
      {"k_eval_group<8, 32, false>", "k_eval_group<16, 64, false>", "k_eval_group<32, 256, false>",
       "k_eval_group<64, 512, false>", "k_eval_group<64, 1024, false>", "k_eval_block<2048, false>",
       "k_eval_block<4096, false>", "hub sequence: k_hubp_plan..k_hubp_finish"},
      {"k_eval_group<8, 32, true>", "k_eval_group<16, 64, true>", "k_eval_group<32, 256, true>",
       "k_eval_group<64, 512, true>", "k_eval_group<64, 1024, true>", "k_eval_block<2048, true>",
       "k_eval_block<4096, true>", "hub sequence: k_hubp_plan..k_hubp_finish"}};
  const float ms = c.prof.stop(s);   // end event right after the launch
  uint32_t before[16];
  std::memcpy(before, c.h_cnt, sizeof(before));
  fold_counters(c, s);
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  c.sync();
  auto u64 = [](const uint32_t *h, int i) { return (double)((uint64_t)h[i] | ((uint64_t)h[i + 1] << 32)); };
  const double visits = u64(c.h_cnt, cVisitsLo) - u64(before, cVisitsLo);
  const double cands = u64(c.h_cnt, cCandsLo) - u64(before, cCandsLo);
  const double units = (double)c.h_cnt[cUnits] - (double)before[cUnits];
  const double wbytes = 8.0;   // fixed-point arc weights at every level
  KernelStat &k = c.prof.get(names[c.precision == 1 ? 1 : 0][bucket]);
  k.launches += 1;
  k.ms += ms;
  k.bytes += visits * (8.0 + wbytes) + units * 48.0 + cands * 32.0;
}

// Algorithmic bytes of the hub partition kernels of one batch, from its hub
// arc visits, hubs and candidate modules (counted on the device) and its arc
// chunks and partitions:
//   plan    per hub: list entry, unit, bounds, three prefix entries (40)
//   owners  per chunk and partition its hub (4), per hub its prefixes (16)
//   count   per arc: neighbour id, its module (gathered), module list entry (12)
//   scatter per arc: module list entry, weight, partitioned key and fixed-point flow (16 + w)
//   eval    per arc: key and flow (12); per candidate its statistics (32; its one-unit bit is cached);
//           per partition: bounds, the unit's own reads and its best (8 + 48 + 28)
//   finish  per partition its best (12), per hub its own reads and flows (64)
static void profile_hubs(Ctx &c, const HubParts &H, const float *tk, bool f32, hipStream_t s) {
  uint32_t before[16];
  std::memcpy(before, c.h_cnt, sizeof(before));
  fold_counters(c, s);
  uint32_t tot[3] = {0, 0, 0};
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  RMX_HIP(hipMemcpyAsync(tot, H.tot, sizeof(tot), hipMemcpyDeviceToHost, s));
  c.sync();
  auto u64 = [](const uint32_t *h, int i) { return (double)((uint64_t)h[i] | ((uint64_t)h[i + 1] << 32)); };
  const double arcs = u64(c.h_cnt, cVisitsLo) - u64(before, cVisitsLo);
  const double cands = u64(c.h_cnt, cCandsLo) - u64(before, cCandsLo);
  const double hubs = (double)c.h_cnt[cUnits] - (double)before[cUnits];
  const double chunks = tot[0], parts = tot[1];
  const double w = 8.0;   // fixed-point arc weight
  c.prof.add("k_hubp_plan", tk[0], hubs * 40.0);
  c.prof.add("k_hubp_owners", tk[1], (chunks + parts) * 4.0 + hubs * 16.0);
  c.prof.add("k_hubp_count", tk[2], arcs * 12.0);
  c.prof.add("k_hubp_scatter", tk[3], arcs * (16.0 + w));
  c.prof.add(f32 ? "k_hubp_eval<true>" : "k_hubp_eval<false>", tk[4], arcs * 12.0 + cands * 32.0 + parts * 84.0);
  c.prof.add("k_hubp_finish", tk[5], parts * 12.0 + hubs * 64.0);
}

// k_commit's and k_commit_big's algorithmic bytes of one batch:
//   per mover: id, CSR bounds, stamp, old and new module, its flows to
//   them (32), its statistics (32) and member count, and the read-modify-
//   write of both module rows and member counts (160): 260 B;
//   per out-arc: the head and its stamp (8); per arc, when the marks are
//   pushed: the neighbour id and its mark (5); per mover with pulled marks
//   its flag (1).
struct CommitProfile {
  float ms = 0;
  unsigned long long arcs[4] = {0, 0, 0, 0};   // k_mover_arcs: small out / all, big out / all
};
static void profile_commit_counts(Ctx &c, Level &L, uint32_t big_deg, CommitProfile &cp, hipStream_t s) {
  DevBuf<unsigned long long> out(c.cache);
  out.alloc(4);
  fill_i64((long long *)out.p, 0, 4, s);
  k_mover_arcs<<<1024, 256, 0, s>>>(c.movers.p, c.counters.p, L.out_ptr.p, L.in_ptr.p, big_deg, out.p);
  RMX_LAUNCH_CHECK();
  RMX_HIP(hipMemcpyAsync(cp.arcs, out.p, sizeof(cp.arcs), hipMemcpyDeviceToHost, s));
  c.sync();
}

void ensure_sweep_buffers(Ctx &c) {
  const int32_t n = c.N;
  if (c.target.n >= n) return;
  c.act.alloc(n);
  c.act_next.alloc(n);
  c.wmark.alloc(n);
  c.target.alloc(n);
  c.newmod.alloc(n);
  c.stamp.alloc(n);
  c.movers.alloc(n);
  c.mv_cap = (uint32_t)((n + 63) / 64 + 1024);
  c.movers_st.alloc((int64_t)kStripes * c.mv_cap);
  c.empty_list.alloc(n);
  c.want_empty.alloc(want_words(n));
  c.rank_tmp.alloc(n);
  c.counters.alloc(kCounterWords);
  c.agg_buf.alloc(4 * (int64_t)n);
  c.batch_keys.alloc(n);
  c.batch_list.alloc(n);
  fill_u8(c.act.p, 1, n, c.stream);
  fill_u8(c.act_next.p, 0, n, c.stream);
  fill_u8(c.wmark.p, 0, n, c.stream);
  fill_u32(c.stamp.p, 0u, n, c.stream);
  fill_u32(c.want_empty.p, 0u, want_words(n), c.stream);
  fill_u32(c.counters.p, 0u, kCounterWords, c.stream);
}

void reset_active(Ctx &c) {
  ensure_sweep_buffers(c);
  fill_u8(c.act.p, 1, c.N, c.stream);
  fill_u8(c.act_next.p, 0, c.N, c.stream);
  fill_u8(c.wmark.p, 0, c.N, c.stream);
}

// The launches of one sweep captured into a HIP graph (when on): capture
// starts at construction; launch() ends it, instantiates and launches the
// graph; the destructor ends an unfinished capture (an exception in between)
// and frees the graph, after the caller has waited for the stream.
struct SweepGraph {
  hipStream_t s;
  bool on, capturing = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  SweepGraph(hipStream_t s_, bool on_) : s(s_), on(on_) {
    if (on) {
      RMX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      capturing = true;
    }
  }
  void launch() {
    if (!on) return;
    capturing = false;
    RMX_HIP(hipStreamEndCapture(s, &g));
    RMX_HIP(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    RMX_HIP(hipGraphLaunch(x, s));
  }
  ~SweepGraph() {
    if (capturing) {
      hipGraph_t t = nullptr;
      (void)hipStreamEndCapture(s, &t);
      if (t) (void)hipGraphDestroy(t);
    }
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
  }
  SweepGraph(const SweepGraph &) = delete;
  SweepGraph &operator=(const SweepGraph &) = delete;
};

// Hub buffers of level L: the hash-table pool and the partitioned
// evaluation's buffers, sized once for all hubs of the level.
static void ensure_hub_buffers(Ctx &c, Level &L, int32_t n_hubs) {
  if (n_hubs > 0 && c.hub_keys.n < L.hub_pool) {
    c.hub_keys.alloc(L.hub_pool);
    c.hub_out.alloc(L.hub_pool);
    c.hub_in.alloc(L.hub_pool);
  }
  if (n_hubs > 0) {
    const int64_t chunks = L.hub_pool / kHubChunk + n_hubs + 1;
    if (c.hub_bd.n < chunks) {
      c.hub_bd.alloc(chunks);
      c.hub_bm.alloc(chunks);
    }
    if (c.hub_plan.n < 3 * ((int64_t)n_hubs + 1)) c.hub_plan.alloc(3 * ((int64_t)n_hubs + 1));
    if (c.big_list.n < n_hubs) c.big_list.alloc(n_hubs);
    if (c.big_plan.n < (int64_t)n_hubs + 1) c.big_plan.alloc((int64_t)n_hubs + 1);
    // partitioned hub evaluation: sized for all hubs of the level at once
    if (c.hp_mtmp.n < L.hub_arcs) {
      c.hp_mtmp.alloc(L.hub_arcs);
      c.hp_ekey.alloc(L.hub_arcs);
      c.hp_ew.alloc(L.hub_arcs);
    }
    if (c.hp_bcnt.n < L.hub_parts + 1) {
      c.hp_bcnt.alloc(L.hub_parts + 1);
      c.hp_bcur.alloc(L.hub_parts + 1);
      c.hp_boff.alloc(L.hub_parts + 1);
      c.hp_bm.alloc(L.hub_parts);
      c.hp_bd.alloc(L.hub_parts);
      c.hp_bfo.alloc(L.hub_parts);
      c.hp_bfi.alloc(L.hub_parts);
    }
    if (c.hp_plan.n < 3 * ((int64_t)n_hubs + 1) + 4) c.hp_plan.alloc(3 * ((int64_t)n_hubs + 1) + 4);
    if (c.hp_phub.n < L.hub_parts) c.hp_phub.alloc(L.hub_parts);
    if (c.hp_chub.n < L.hub_arcs / kHubArcChunk + n_hubs + 1) c.hp_chub.alloc(L.hub_arcs / kHubArcChunk + n_hubs + 1);
    if (c.hp_ovf.n < n_hubs) {
      c.hp_ovf.alloc(n_hubs);
      c.hp_own.alloc(2 * (int64_t)n_hubs);
      c.hp_hubu.alloc(n_hubs);
    }
  }
}

// The empty modules available to a sweep (ascending id; none in a
// sub-module partition: its stale member lists never exceed one member, so
// the reference never offers the option there), and the sweep's units per
// (bucket, batch), inactive units last: one stable 8-bit radix pass over the
// bucket positions.  seg[k]..seg[k+1] is the range of key
// k = bucket * (B + 1) + batch in c.batch_list.
static void sweep_lists(Ctx &c, Level &L, bool prioritized, uint64_t sweep_seed, bool batch, int B,
                        std::vector<uint32_t> &seg, uint32_t &n_empty) {
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  n_empty = 0;
  if (!c.sub_mode) {
    k_empty_flags<<<grid_for(c.N, 256), 256, 0, s>>>(c.mNodes.p, c.N, c.rank_tmp.p);
    RMX_LAUNCH_CHECK();
    exclusive_scan_u32(c.rank_tmp.p, c.target.p, c.N, c.counters.p + 8, c.scratch, s);
    k_compact<<<grid_for(c.N, 256), 256, 0, s>>>(c.rank_tmp.p, c.target.p, c.N, c.empty_list.p);
    RMX_LAUNCH_CHECK();
  }
  const int NK = kNumBuckets * (B + 1);
  seg.assign(NK + 1, 0);
  {
    const int *o = L.bucket_off;
    k_batch_keys<<<grid_for(n, 256), 256, 0, s>>>(L.bucket.p, n, make_int4(o[0], o[1], o[2], o[3]),
                                                  make_int4(o[4], o[5], o[6], o[7]), c.act.p, prioritized ? 0 : 1,
                                                  sweep_seed, batch ? L.uinst.p : nullptr, c.b_seed.p, B,
                                                  c.batch_keys.p, c.batch_list.p);
    RMX_LAUNCH_CHECK();
    radix_sort_pairs_u32(c.batch_keys.p, c.batch_list.p, n, 0, 8, c.scratch2, s);
    fill_u32(c.counters.p + 16, 0u, 256, s);
    k_hist256<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, s>>>(c.batch_keys.p, n,
                                                                               c.counters.p + 16);
    RMX_LAUNCH_CHECK();
    // one wait for the host: the empty-module count (word 8) and the
    // (bucket, batch) histogram (words 16..271)
    std::vector<uint32_t> h(16 + 256);
    RMX_HIP(hipMemcpyAsync(h.data(), c.counters.p, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, s));
    c.sync();
    for (int k = 0; k < NK; ++k) seg[k + 1] = seg[k] + h[16 + k];
    n_empty = c.sub_mode ? 0u : h[8];
  }
}

// The global-table evaluation of a batch's hubs whose partition table
// overflowed (its plan gives the other hubs no work).
static void hub_fallback(Ctx &c, EvalArgs a, const uint32_t *list, int32_t cnt, bool f32, hipStream_t s) {
  a.list = list;
  a.list_n = cnt;
  HubPlan P{c.hub_plan.p, c.hub_plan.p + (cnt + 1), c.hub_plan.p + 2 * (cnt + 1), c.hub_bd.p, c.hub_bm.p,
            c.hp_ovf.p};
  k_hub_plan<<<1, 256, 0, s>>>(a, P);
  k_hub_clear<<<4096, 256, 0, s>>>(a, P);
  k_hub_insert<<<4096, 256, 0, s>>>(a, P);
  if (f32) k_hub_eval<true><<<4096, 256, 0, s>>>(a, P);
  else k_hub_eval<false><<<4096, 256, 0, s>>>(a, P);
  k_hub_finish<<<(unsigned)std::min<int64_t>(1024, (cnt + 3) / 4), 256, 0, s>>>(a, P);
  RMX_LAUNCH_CHECK();
  fold_counters(c, s);
}

// The evaluation launches of batch b of a sweep (every degree bucket, hubs
// included); returns the units evaluated, and the batch's hubs for the
// overflow fallback in hub_list / hub_n.
static uint32_t eval_batch(Ctx &c, Level &L, EvalArgs &a, const std::vector<uint32_t> &seg, int B, int b,
                           bool f32, hipStream_t s, const uint32_t *&hub_list, int32_t &hub_n) {
  hub_list = nullptr;
  hub_n = 0;
  uint32_t units_b = 0;   // units evaluated in this batch
  for (int bk = 0; bk < kNumBuckets; ++bk) {
    const int key = bk * (B + 1) + b;
    const int32_t cnt = (int32_t)(seg[key + 1] - seg[key]);
    if (cnt <= 0) continue;
    units_b += (uint32_t)cnt;
    a.list = c.batch_list.p + seg[key];
    a.list_n = cnt;
    bool hub_profiled = false;   // the partitioned hub kernels account themselves
    a.bound = a.smax ? c.bound_mode[bk] : 0;
    if (c.prof.on) c.prof.begin(s);
    switch (bk) {
      case 0: case 1: case 2: case 3: case 4: case 5: case 6:
        launch_bucket(bk, a, cnt, f32, s);
        break;
      default: {
        if (c.hub_global == 1) {
          HubPlan P{c.hub_plan.p, c.hub_plan.p + (cnt + 1), c.hub_plan.p + 2 * (cnt + 1), c.hub_bd.p,
                    c.hub_bm.p, nullptr};
          k_hub_plan<<<1, 256, 0, s>>>(a, P);
          k_hub_clear<<<4096, 256, 0, s>>>(a, P);
          k_hub_insert<<<4096, 256, 0, s>>>(a, P);
          if (f32) k_hub_eval<true><<<4096, 256, 0, s>>>(a, P);
          else k_hub_eval<false><<<4096, 256, 0, s>>>(a, P);
          k_hub_finish<<<(unsigned)std::min<int64_t>(1024, (cnt + 3) / 4), 256, 0, s>>>(a, P);
          break;
        }
        const int64_t NB = L.hub_parts;
        HubParts H;
        H.A = c.hp_plan.p;
        H.PB = H.A + (cnt + 1);
        H.AB = H.PB + (cnt + 1);
        H.tot = H.AB + (cnt + 1);
        H.bcnt = c.hp_bcnt.p;
        H.bcur = c.hp_bcur.p;
        H.boff = c.hp_boff.p;
        H.mtmp = c.hp_mtmp.p;
        H.ekey = c.hp_ekey.p;
        H.ew = c.hp_ew.p;
        H.own = c.hp_own.p;
        H.ovf = c.hp_ovf.p;
        H.bd = c.hp_bd.p;
        H.bm = c.hp_bm.p;
        H.bfo = c.hp_bfo.p;
        H.bfi = c.hp_bfi.p;
        H.force_ovf = c.hub_global == 2 ? 1 : 0;
        H.hubu = c.hp_hubu.p;
        H.chub = c.hp_chub.p;
        H.phub = c.hp_phub.p;
        // profiling: each kernel of the sequence timed on its own
        float tk[6] = {0, 0, 0, 0, 0, 0};
        auto t0 = [&]() { if (c.prof.on) c.prof.begin(s); };
        auto t1 = [&](int k) { if (c.prof.on) tk[k] = c.prof.stop(s); };
        t0();
        k_hubp_plan<<<1, kPlanBlock, 0, s>>>(a, H);
        t1(0);
        t0();
        k_hubp_owners<<<grid_for((int64_t)cnt * kWave, 256), 256, 0, s>>>(a, H);
        t1(1);
        t0();
        k_hubp_count<<<4096, 256, 0, s>>>(a, H);
        t1(2);
        // partitions beyond this batch's hold stale counts: only boff[0..tot[1]] is read
        exclusive_scan_u32(H.bcnt, H.boff, NB + 1, nullptr, c.scratch, s);
        t0();
        k_hubp_scatter<<<4096, 256, 0, s>>>(a, H);
        t1(3);
        t0();
        if (f32) k_hubp_eval<true><<<8192, 256, 0, s>>>(a, H);
        else k_hubp_eval<false><<<8192, 256, 0, s>>>(a, H);
        t1(4);
        t0();
        k_hubp_finish<<<(unsigned)std::min<int64_t>(1024, (cnt + 3) / 4), 256, 0, s>>>(a, H);
        t1(5);
        RMX_LAUNCH_CHECK();
        if (c.prof.on) {
          profile_hubs(c, H, tk, f32, s);
          hub_profiled = true;
        }
        hub_list = a.list;
        hub_n = cnt;
        break;
      }
    }
    RMX_LAUNCH_CHECK();
    if (c.prof.on && !hub_profiled) profile_eval(c, bk, s);
  }
  return units_b;
}

int64_t move_sweep(Ctx &c, bool prioritized, double vthresh, uint64_t seed, int64_t *edge_visits) {
  ensure_sweep_buffers(c);
  Level &L = c.cur();
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  const uint64_t sweep_seed = mix64(seed ^ mix64(++c.sweep_counter));
  const bool batch = c.nb > 0 && !c.sub_mode;   // independent instances (not a sub-module partition)

  fill_u32(c.counters.p + 2, 0u, 6, s);   // moved, visits, candidates
  fill_u32(c.counters.p + cUnits, 0u, 1, s);
  fill_u32(c.counters.p + cErr, 0u, 2, s);   // error, pull flag
  fill_u32(c.counters.p + kStripeBase, 0u, kStripes * kStripeWords, s);
  const int32_t n_hubs = L.bucket_off[kNumBuckets] - L.bucket_off[kNumBuckets - 1];
  ensure_hub_buffers(c, L, n_hubs);
  const int B = sweep_batches(c, L);
  std::vector<uint32_t> seg;
  uint32_t n_empty = 0;
  sweep_lists(c, L, prioritized, sweep_seed, batch, B, seg, n_empty);
  uint32_t empty_cursor = 0;
  if (batch) {   // each instance's empty modules
    c.b_ebeg.alloc(c.nb);
    c.b_ecnt.alloc(c.nb);
    c.b_ecur.alloc(c.nb);
    k_bempty_seg<<<grid_for(c.nb, 256), 256, 0, s>>>(c.nb, c.b_off.p, c.empty_list.p, n_empty, c.b_ebeg.p,
                                                      c.b_ecnt.p, c.b_ecur.p);
    RMX_LAUNCH_CHECK();
  }
  if (c.prof.on) {   // h_cnt mirrors the counters from here on (profiling deltas)
    RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
    c.sync();
  }

  EvalArgs a = sweep_args(c, L, prioritized, vthresh, sweep_seed, batch);
  // one network: S after each batch from the modules' exit probabilities,
  // refreshed here and then for the modules each batch touches
  const bool single = !batch && !c.sub_mode;
  if (single) exit_probs_refresh(c);

  ArcW aout = arcw_out(c, L);
  const bool f32 = c.precision == 1;   // mixed precision: candidates in FP32

  // A batch of instances or a sub-module partition keeps its empty-module
  // cursors on the device (or offers none), so its sweep needs no host wait:
  // the resolve / commit kernels read the mover count on the device and run
  // on grids sized for every unit of the batch, the want-empty scan runs in
  // every batch, and the hub fallback is launched whenever hubs were
  // evaluated.  The sweep's launches are then captured into one graph.
  const bool quiet = c.use_graph && !c.prof.on && (batch || c.sub_mode);
  if (quiet) {   // nothing may allocate while the stream is captured
    const int64_t nb1 = n_hubs > 0 ? L.hub_parts + 1 : 0;
    exclusive_scan_reserve(std::max<int64_t>(n, nb1), c.scratch);
    if (batch) {
      c.b_S.alloc(c.nb);
      c.b_L.alloc(c.nb);
      c.b_nmod.alloc(c.nb);
    }
  }
  SweepGraph graph(s, quiet);
  bool maybe_pull = false;   // some batch may have left its marks to k_pull_marks
  for (int b = 0; b < B; ++b) {
    const uint32_t *hub_list = nullptr;   // this batch's hubs, for the overflow fallback
    int32_t hub_n = 0;
    a.batch = b;
    a.empty_avail = empty_cursor < n_empty ? 1 : 0;
    k_reset_counters<<<1, 64, 0, s>>>(c.counters.p);
    RMX_LAUNCH_CHECK();
    const uint32_t units_b = eval_batch(c, L, a, seg, B, b, f32, s, hub_list, hub_n);
    fold_counters(c, s);
    if (units_b == 0) continue;
    uint32_t n_movers = units_b, n_want = 1;   // bounds when not read back
    if (quiet) {
      if (hub_n > 0) hub_fallback(c, a, hub_list, hub_n, f32, s);
    } else {
      RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
      c.sync();
      if (c.h_cnt[cErr]) throw HipError("move sweep: mover list of a counter copy overflowed");
      if (hub_n > 0 && c.h_cnt[cHubOvf] > 0) {   // a partition table overflowed: those hubs again
        hub_fallback(c, a, hub_list, hub_n, f32, s);
        RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
        c.sync();
        if (c.h_cnt[cErr]) throw HipError("move sweep: mover list of a counter copy overflowed");
      }
      n_movers = c.h_cnt[cMovers];
      n_want = c.h_cnt[cWant];
      if (n_movers == 0) continue;
    }
    ++c.stamp_now;
    if (batch) {
      if (n_want > 0) want_prefix(c, n, c.counters.p + 9, s);
      k_resolve_batch<<<stride_grid(n_movers, 256), 256, 0, s>>>(
          c.movers.p, c.counters.p, c.target.p, c.want_empty.p, c.rank_tmp.p, c.counters.p + 9, n, L.uinst.p,
          L.uoff.p, c.empty_list.p, c.b_ebeg.p, c.b_ecnt.p, c.b_ecur.p, a.empty_key, c.newmod.p, c.stamp.p,
          c.stamp_now);
      RMX_LAUNCH_CHECK();
      if (n_want > 0) {
        k_bempty_advance<<<grid_for(c.nb, 256), 256, 0, s>>>(c.nb, L.uoff.p, c.want_empty.p, c.rank_tmp.p,
                                                              c.counters.p + 9, n, c.b_ecnt.p, c.b_ecur.p);
        RMX_LAUNCH_CHECK();
        fill_u32(c.want_empty.p, 0u, want_words(n), s);   // the requests are served
      }
    } else {
      if (n_want > 0 && n_empty > 0) want_prefix(c, n, nullptr, s);
      k_resolve<<<stride_grid(n_movers, 256), 256, 0, s>>>(c.movers.p, c.counters.p, c.target.p, c.want_empty.p,
                                                            c.rank_tmp.p, c.empty_list.p, empty_cursor, n_empty,
                                                            a.empty_key, c.newmod.p, c.stamp.p, c.stamp_now);
      RMX_LAUNCH_CHECK();
      if (n_want > 0) {
        fill_u32(c.want_empty.p, 0u, want_words(n), s);   // the requests are served
        uint32_t left = n_empty - empty_cursor;
        empty_cursor += n_want < left ? n_want : left;
      }
    }
    const uint32_t big_deg = n_hubs > 0 ? (uint32_t)c.cuts.v[kNumBuckets - 2] : 0xffffffffu;
    // marks by pull when at least half of a large batch moves
    const uint32_t pull_min = prioritized && units_b >= kPullMinUnits ? (units_b + 1) / 2 : 0xffffffffu;
    if (pull_min != 0xffffffffu && (quiet || n_movers >= pull_min)) maybe_pull = true;
    CommitProfile cprof;
    if (c.prof.on) c.prof.begin(s);
    k_commit<16><<<stride_grid((int64_t)n_movers * 16, 256), 256, 0, s>>>(
        c.movers.p, c.counters.p, L.out_ptr.p, L.out_idx.p, L.in_ptr.p, L.in_idx.p, aout, L.mod.p, c.newmod.p,
        c.stamp.p, c.stamp_now, c.agg_buf.p, L.us.p, L.nmem.p, c.ms.p, c.mNodes.p, c.mUnits.p, c.act_next.p,
        prioritized ? 1 : 0, c.counters.p + kStripeBase, big_deg, c.big_list.p, c.counters.p + cBig, pull_min,
        c.wmark.p, c.counters.p + cPull, c.smax.p);
    RMX_LAUNCH_CHECK();
    const bool pushed_marks = prioritized && n_movers < pull_min;
    if (c.prof.on) {
      cprof.ms = c.prof.stop(s);
      profile_commit_counts(c, L, big_deg, cprof, s);
      const double pulled = (prioritized && !pushed_marks) ? (double)n_movers : 0.0;
      c.prof.add("k_commit<16>", cprof.ms,
                 (double)n_movers * 260.0 + (double)cprof.arcs[0] * 8.0 +
                     (pushed_marks ? (double)cprof.arcs[1] * 5.0 : 0.0) + pulled);
    }
    if (n_hubs > 0) {
      // only hubs (more than big_deg arcs) are deferred: the list holds at most n_hubs
      k_commit_plan<<<1, 256, 0, s>>>(c.big_list.p, c.counters.p + cBig, L.out_ptr.p, L.in_ptr.p, c.big_plan.p);
      RMX_LAUNCH_CHECK();
      RMX_PROF_BEGIN(c);
      k_commit_big<<<2048, 256, 0, s>>>(c.big_list.p, c.counters.p + cBig, c.big_plan.p, L.out_ptr.p, L.out_idx.p,
                                        L.in_ptr.p, L.in_idx.p, aout, L.mod.p, c.newmod.p, c.stamp.p, c.stamp_now,
                                        c.ms.p, c.act_next.p, prioritized ? 1 : 0, c.counters.p, pull_min);
      RMX_LAUNCH_CHECK();
      RMX_PROF_END(c, "k_commit_big", (double)cprof.arcs[2] * 8.0 + (pushed_marks ? (double)cprof.arcs[3] * 5.0 : 0.0));
    }
    if (single) {
      k_qv_touched<<<stride_grid(n_movers, 256), 256, 0, s>>>(c.movers.p, c.counters.p, c.stamp.p, c.stamp_now,
                                                               L.mod.p, c.newmod.p, c.ms.p, c.mNodes.p, c.qv.p);
      RMX_LAUNCH_CHECK();
    }
    k_apply<<<stride_grid(n_movers, 256), 256, 0, s>>>(c.movers.p, c.counters.p, c.newmod.p, c.stamp.p,
                                                        c.stamp_now, L.mod.p, c.mUnits.p, c.mOne.p);
    RMX_LAUNCH_CHECK();
    // S for the next batch
    if (c.sub_mode) instance_totals_async(c);   // of every instance
    else if (batch) module_totals_async(c);     // of every instance
    else sum_exit_async(c);
  }
  graph.launch();
  if (prioritized) {
    if (maybe_pull) {
      k_pull_marks<8><<<stride_grid((int64_t)n * 8, 256), 256, 0, s>>>(
          c.counters.p + cPull, n, L.out_ptr.p, L.out_idx.p, L.in_ptr.p, L.in_idx.p, c.wmark.p, c.act_next.p);
      RMX_LAUNCH_CHECK();
      fill_u8(c.wmark.p, 0, n, s);
    }
    c.act.swap(c.act_next);
    fill_u8(c.act_next.p, 0, c.N, s);
  }
  fold_counters(c, s);   // moved nodes of the last batch
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  if (c.sub_mode) {
    instance_totals_async(c);
    c.sync();
  } else {
    module_totals(c, &c.sum_exit, &c.codelength);
  }
  if (c.h_cnt[cErr]) throw HipError("move sweep: mover list of a counter copy overflowed");
  const uint32_t moved = c.h_cnt[cMoved];
  unsigned long long vis = (unsigned long long)c.h_cnt[cVisitsLo] | ((unsigned long long)c.h_cnt[cVisitsLo + 1] << 32);
  if (edge_visits) *edge_visits = (int64_t)vis;
  c.timing.edge_visits += (int64_t)vis;
  c.timing.sweeps++;
  return moved;
}

void evaluate_moves(Ctx &c, uint32_t *d_best_m, double *d_best_d) {
  ensure_sweep_buffers(c);
  Level &L = c.cur();
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  fill_u32(c.counters.p + 2, 0u, 6, s);
  fill_u32(c.counters.p + cUnits, 0u, 1, s);
  fill_u32(c.counters.p + cErr, 0u, 2, s);
  fill_u32(c.counters.p + kStripeBase, 0u, kStripes * kStripeWords, s);
  const int32_t n_hubs = L.bucket_off[kNumBuckets] - L.bucket_off[kNumBuckets - 1];
  ensure_hub_buffers(c, L, n_hubs);
  std::vector<uint32_t> seg;
  uint32_t n_empty = 0;
  sweep_lists(c, L, false, 0, false, 1, seg, n_empty);   // every unit, one batch
  module_totals_async(c);                                 // S of the current partition in c.scal
  EvalArgs a = sweep_args(c, L, false, 0.0, 0, false);
  a.batches = 1;
  a.batch = 0;
  a.vthresh = -HUGE_VAL;   // records every unit's best and emits no mover
  a.empty_avail = n_empty > 0 ? 1 : 0;
  a.best_m = d_best_m;
  a.best_d = d_best_d;
  fill_u32(d_best_m, kNone, n, s);
  k_reset_counters<<<1, 64, 0, s>>>(c.counters.p);
  RMX_LAUNCH_CHECK();
  const bool f32 = c.precision == 1;
  const uint32_t *hub_list = nullptr;
  int32_t hub_n = 0;
  eval_batch(c, L, a, seg, 1, 0, f32, s, hub_list, hub_n);
  fold_counters(c, s);
  RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
  c.sync();
  if (hub_n > 0 && c.h_cnt[cHubOvf] > 0) {   // a partition table overflowed: those hubs again
    hub_fallback(c, a, hub_list, hub_n, f32, s);
    RMX_HIP(hipMemcpyAsync(c.h_cnt, c.counters.p, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, s));
    c.sync();
  }
  if (c.h_cnt[cErr]) throw HipError("evaluation: an error flag was raised");
}

int64_t n_active(Ctx &c) {
  std::vector<uint8_t> h(c.N);
  RMX_HIP(hipMemcpyAsync(h.data(), c.act.p, c.N, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  int64_t k = 0;
  for (int32_t i = 0; i < c.cur().n; ++i) k += h[i];
  return k;
}

BucketCuts autotune_buckets(Ctx &c) {
  ensure_sweep_buffers(c);
  Level &L = c.cur();
  hipStream_t s = c.stream;
  const int32_t n = L.n;
  BucketCuts cuts = c.cuts;
  if (n <= 0) return cuts;
  // the level's units grouped by degree class
  DevBuf<uint32_t> key(c.cache), list(c.cache), cnt(c.cache);
  key.alloc(n);
  list.alloc(n);
  cnt.alloc(256);
  k_class_key<<<grid_for(n, 256), 256, 0, s>>>(L.out_ptr.p, L.in_ptr.p, n, key.p, list.p);
  RMX_LAUNCH_CHECK();
  radix_sort_pairs_u32(key.p, list.p, n, 0, 8, c.scratch, s);
  fill_u32(cnt.p, 0u, 256, s);
  k_hist256<<<(unsigned)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, s>>>(key.p, n, cnt.p);
  RMX_LAUNCH_CHECK();
  std::vector<uint32_t> h(256);
  RMX_HIP(hipMemcpyAsync(h.data(), cnt.p, sizeof(uint32_t) * 256, hipMemcpyDeviceToHost, s));
  c.sync();
  std::vector<uint32_t> off(kNumClasses + 1, 0);
  for (int k = 0; k < kNumClasses; ++k) off[k + 1] = off[k] + h[k];

  // evaluation only (no empty-module option, nothing committed)
  const bool batch = c.nb > 0 && !c.sub_mode;
  EvalArgs a = sweep_args(c, L, false, 0.0, 0, batch);
  a.empty_avail = 0;
  a.e_cnt = nullptr;
  a.e_cur = nullptr;
  a.batch = 0;
  const bool f32 = c.precision == 1;
  constexpr int NK = kNumBuckets - 1;   // evaluation kernels below the hubs
  constexpr double kInf = 1e300;
  constexpr uint32_t kMinUnits = 4096;  // smaller classes keep any eligible kernel
  double cost[kNumClasses][NK];
  hipEvent_t e0, e1;
  RMX_HIP(hipEventCreate(&e0));
  RMX_HIP(hipEventCreate(&e1));
  for (int cl = 0; cl < kNumClasses; ++cl) {
    const uint32_t units = h[cl];
    int natural = 0;   // the smallest kernel whose table fits the class
    while (natural < NK && kBucketMaxDeg[natural] < kClassMax[cl]) ++natural;
    for (int k = 0; k < NK; ++k) {
      cost[cl][k] = kInf;
      if (kBucketMaxDeg[k] < kClassMax[cl]) continue;        // table too small
      if (units < kMinUnits) { cost[cl][k] = 0.0; continue; }
      if (k > natural + 2) continue;                         // far too large tables
      a.list = list.p + off[cl];
      a.list_n = (int32_t)units;
      float best = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        k_reset_counters<<<1, 64, 0, s>>>(c.counters.p);
        fill_u32(c.counters.p + kStripeBase, 0u, kStripes * kStripeWords, s);
        RMX_HIP(hipEventRecord(e0, s));
        launch_bucket(k, a, (int32_t)units, f32, s);
        RMX_LAUNCH_CHECK();
        RMX_HIP(hipEventRecord(e1, s));
        RMX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RMX_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = ms < best ? ms : best;
      }
      cost[cl][k] = best;
    }
  }
  RMX_HIP(hipEventDestroy(e0));
  RMX_HIP(hipEventDestroy(e1));
  k_reset_counters<<<1, 64, 0, s>>>(c.counters.p);
  fill_u32(c.counters.p + kStripeBase, 0u, kStripes * kStripeWords, s);
  fill_u32(c.counters.p, 0u, 16, s);
  c.sync();

  // cheapest assignment of classes to kernels that keeps kernel order
  // non-decreasing with degree (so that it is a set of cut-offs)
  double best[kNumClasses][NK];
  int from[kNumClasses][NK];
  for (int cl = 0; cl < kNumClasses; ++cl)
    for (int k = 0; k < NK; ++k) {
      best[cl][k] = kInf;
      from[cl][k] = -1;
      if (cost[cl][k] >= kInf) continue;
      if (cl == 0) {
        best[cl][k] = cost[cl][k];
        continue;
      }
      for (int j = 0; j <= k; ++j)
        if (best[cl - 1][j] < kInf && best[cl - 1][j] + cost[cl][k] < best[cl][k]) {
          best[cl][k] = best[cl - 1][j] + cost[cl][k];
          from[cl][k] = j;
        }
    }
  int k_last = 0;
  for (int k = 1; k < NK; ++k)
    if (best[kNumClasses - 1][k] < best[kNumClasses - 1][k_last]) k_last = k;
  if (best[kNumClasses - 1][k_last] >= kInf) return cuts;   // cannot happen: the default is feasible
  int asg[kNumClasses];
  for (int cl = kNumClasses - 1, k = k_last; cl >= 0; --cl) {
    asg[cl] = k;
    k = from[cl][k];
  }
  for (int k = 0; k < NK; ++k) {
    int32_t v = k == 0 ? 0 : cuts.v[k - 1];
    for (int cl = 0; cl < kNumClasses; ++cl)
      if (asg[cl] == k) v = kClassMax[cl];
    cuts.v[k] = v;
  }
  cuts.v[NK] = kBucketMaxDeg[NK];
  return cuts;
}

}  // namespace rmx
