// Device-resident state of one network (one HIP stream, one GPU).
#pragma once

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "common.h"
#include "prims.h"

namespace rmx {

// Degree buckets of the move sweep: units are evaluated by groups of G
// lanes whose LDS hash table holds H keys; the last bucket (hubs) uses one
// workgroup per unit with its table in global memory.
constexpr int kNumBuckets = 8;
// upper bound of (out + in) degree per bucket; a table holds at least twice
// the degree, so the cut-offs follow the table sizes (and with them the LDS
// per workgroup, which sets how many workgroups share a CU)
constexpr int kBucketMaxDeg[kNumBuckets] = {16, 32, 128, 256, 512, 1023, 2048, 0x7fffffff};
// Upper degree bounds of the buckets in use (the defaults above, or the
// autotuner's): bucket b takes the units with degree in (v[b-1], v[b]]; a
// bound never exceeds the default, the table size of the bucket's kernel.
struct BucketCuts {
  int32_t v[kNumBuckets];
};
// batches per sweep: (bucket, batch) sort keys must fit one 8-bit digit
constexpr int kMaxBatches = 256 / kNumBuckets - 1;
// Hubs are evaluated in partitions of about kHubPart arcs (split by a hash of
// the neighbour's module), each aggregated in an LDS table.
constexpr int kHubPart = 1024;
constexpr int kTotChunk = 4096;   // module ids per workgroup of the per-instance totals
// Loads that a lane of a G-lane row kernel keeps in flight: a hub row of 60k
// arcs then costs its 16 lanes about 470 round trips to memory instead of
// 3750.  Each lane still adds its arcs in the order of the plain loop (k, k +
// G, k + 2G, ...), so every sum is bit-identical to it.
constexpr int kRowAhead = 8;

// Device blocks kept for reuse within one context.  All work of a context
// runs on one stream, so a block handed back while kernels that use it are
// still queued can be handed out again at once: later launches are ordered
// after them.  Blocks are released when the context is destroyed.
struct BlockCache {
  std::vector<std::pair<size_t, void *>> free_blocks;
  std::vector<void *> all;
  void *get(size_t bytes, size_t *got) {
    size_t best = (size_t)-1;
    int bi = -1;
    for (size_t i = 0; i < free_blocks.size(); ++i) {
      size_t b = free_blocks[i].first;
      if (b >= bytes && b <= 4 * bytes + (1u << 20) && b < best) {
        best = b;
        bi = (int)i;
      }
    }
    if (bi >= 0) {
      void *p = free_blocks[bi].second;
      *got = free_blocks[bi].first;
      free_blocks.erase(free_blocks.begin() + bi);
      return p;
    }
    size_t b = (bytes + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
    void *p = nullptr;
    RMX_HIP(hipMalloc(&p, b));
    all.push_back(p);
    *got = b;
    return p;
  }
  void put(void *p, size_t bytes) { free_blocks.push_back({bytes, p}); }
  void release() {
    for (void *p : all) (void)hipFree(p);
    all.clear();
    free_blocks.clear();
  }
};

// Owning device allocation.  alloc() keeps the block when it is big enough.
// With a BlockCache the block comes from and returns to the cache; without
// one it is hipMalloc'ed and hipFree'd (hipFree waits for the device).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  int64_t n = 0;
  size_t cap_bytes = 0;
  BlockCache *cache = nullptr;
  DevBuf() = default;
  explicit DevBuf(BlockCache &bc) : cache(&bc) {}
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { free(); }
  void alloc(int64_t count) {
    const size_t need = sizeof(T) * (size_t)(count > 0 ? count : 0);
    if (p && need <= cap_bytes) {
      n = count;
      return;
    }
    free();
    if (count > 0) {
      if (cache) {
        p = (T *)cache->get(need, &cap_bytes);
      } else {
        RMX_HIP(hipMalloc(&p, need));
        cap_bytes = need;
      }
    }
    n = count;
  }
  void free() {
    if (p) {
      if (cache) cache->put(p, cap_bytes);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    cap_bytes = 0;
  }
  T *get() const { return p; }
  void swap(DevBuf &o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap_bytes, o.cap_bytes);
    std::swap(cache, o.cache);
  }
};

// One level of the hierarchy: the original nodes (level 0) or the super
// nodes of the current contraction (level 1).  Units are nodes or super
// nodes; arcs carry flows f_ab (level 0 after the steady state) or summed
// inter-module flows (level 1).
struct Level {
  int32_t n = 0;
  int64_t m = 0;
  DevBuf<uint32_t> out_ptr, out_idx, in_ptr, in_idx;
  DevBuf<double> out_w, in_w;     // FP64 flows
  DevBuf<float> out_wf, in_wf;    // FP32 mirror (mixed precision)
  DevBuf<long long> out_wx, in_wx;  // level 1: exact fixed-point arc flow sums
  DevBuf<double> p, tau;          // level 0: steady-state flow and teleport weight
  // unit statistics in fixed point: flow, teleport, dangling flow, out flow
  DevBuf<Stat4> us;
  DevBuf<int32_t> nmem;           // original nodes inside the unit
  DevBuf<uint32_t> mod;           // module of each unit
  DevBuf<uint32_t> bucket;        // units grouped by degree bucket
  int32_t bucket_off[kNumBuckets + 1] = {0};
  // hub units (last bucket): each owns a global hash table region of
  // nextpow2(2 deg + 2) keys; hub_off[i] is the region of the i-th hub
  DevBuf<long long> hub_off;
  int64_t hub_pool = 0;
  int64_t hub_arcs = 0;    // sum of the hubs' degrees
  int64_t hub_parts = 0;   // partitions of all hubs (partitioned hub evaluation)
  int32_t max_deg = 0;
  // batch of instances: instance of every unit, and the first unit of every
  // instance (units are grouped by instance, in instance order)
  DevBuf<uint32_t> uinst, uoff;
  void free_all();
  void swap(Level &o);
};

struct Timing {
  double flow_ms = 0, sweep_ms = 0, contract_ms = 0;
  int64_t edge_visits = 0;
  int sweeps = 0;
};

// Per-kernel launch statistics of the profiling mode: HIP events on the
// kernel's own stream around each launch, and the launch's algorithmic HBM
// bytes (the bytes the algorithm must move, not the cache traffic).
struct KernelStat {
  std::string name;
  int64_t launches = 0;
  double ms = 0, bytes = 0;
};

// Device view of the chunk plan of large instances (coarse.hip)
struct BigInst {
  const uint32_t *list;    // big instance -> its segment
  const uint32_t *base;    // big instance -> its first chunk slot
  const uint32_t *owner;   // chunk slot -> big instance
  double *part;            // chunk slot -> partial sums
};

struct Profiler {
  bool on = false;
  float last_ms = 0;
  std::vector<KernelStat> stats;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  KernelStat &get(const std::string &name) {
    for (auto &k : stats)
      if (k.name == name) return k;
    stats.push_back(KernelStat{name});
    return stats.back();
  }
  void begin(hipStream_t s);
  // records the end event and waits; returns the launch's milliseconds
  float stop(hipStream_t s);
  // stop() and account the launch under `name`
  void end(hipStream_t s, const std::string &name, double bytes);
  // account a launch already timed with begin() / stop()
  void add(const std::string &name, float ms, double bytes) {
    KernelStat &k = get(name);
    k.launches += 1;
    k.ms += ms;
    k.bytes += bytes;
  }
};

// Profiling mode: one launch bracketed by events on the context's stream
// (the byte count is evaluated only when profiling).
#define RMX_PROF_BEGIN(c)                      \
  do {                                         \
    if ((c).prof.on) (c).prof.begin((c).stream); \
  } while (0)
#define RMX_PROF_END(c, name, bytes)                             \
  do {                                                           \
    if ((c).prof.on) (c).prof.end((c).stream, (name), (double)(bytes)); \
  } while (0)

// The bin-major slot layout of the blocked push / pull steady state
// (flow.hip): bins of R consecutive rows, every out-arc's slot in its bin and
// every slot's head relative to the bin, plus the products' array.  It depends
// only on the network's structure and the product type, so it is kept from
// one steady state to the next (Ctx::pb) until the next upload.
struct PushBins {
  int32_t R = 0, nbins = 0, rshift = 0;
  int64_t ntiles = 0;
  bool fp64 = false;   // FP64 products (val64) instead of FP32 (val)
  DevBuf<uint32_t> bptr, slot, trow;
  DevBuf<uint16_t> bdst;
  // FP32 products staged by tile (flow.hip, k_pb_push_staged): tiles of
  // kStageTile arcs, each arc's position in its tile's bin-major order, the
  // slot of every such position, and the row of every 8th arc
  bool staged = false;
  int64_t stiles = 0;
  DevBuf<uint16_t> lpos;
  DevBuf<uint32_t> gslot, grow;
  DevBuf<float> val;
  DevBuf<double> val64;
  void alloc_values(int64_t m) {
    if (fp64) val64.alloc(m);
    else val.alloc(m);
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  int32_t N = 0;
  int64_t E = 0;
  int precision = 0;   // 0 FP64, 1 mixed (FP32 bulk, FP64 accumulation)
  int batches = 0;     // sub-sweeps per sweep; 0: chosen per level (sweep_batches)
  int hub_global = 0;  // 1: hubs with global-memory tables instead of LDS partitions
  int fp64_blocked = 1;   // FP64 steady state: 1 blocked push / pull, 0 the in-CSR pull
  int use_graph = 0;   // sweeps of a batch or a sub-module partition as one captured graph
  int attempts_sequential = 0;   // 1: attempts one after another even where a batch of copies fits
  // candidate bound of the evaluation kernels per degree bucket (7: the hub
  // partitions): 1 on, 0 off (units of up to 32 arcs have no bound code:
  // one round of candidate reads, nothing to gain)
  int bound_mode[8] = {0, 0, 1, 1, 1, 1, 1, 1};
  BucketCuts cuts = {{kBucketMaxDeg[0], kBucketMaxDeg[1], kBucketMaxDeg[2], kBucketMaxDeg[3], kBucketMaxDeg[4],
                      kBucketMaxDeg[5], kBucketMaxDeg[6], kBucketMaxDeg[7]}};
  bool initiated = false;

  Level L0, L1;
  Level L2;            // the next contraction is built here, then swapped with L1
  int level = 0;
  DevBuf<double> node_w, raw_w;   // inputs kept for re-initiation
  DevBuf<uint8_t> dangling;       // level-0 dangling flag
  DevBuf<double> raw_in;          // level-0: the raw arc weights in in-CSR order
  DevBuf<uint32_t> node_sp;       // node -> super node (level 1)

  // module SoA, N slots (module ids are node ids, as in the reference)
  DevBuf<Stat4> ms;                     // fixed-point sums: flow, teleport, dangling, exit arcs
  DevBuf<unsigned long long> smax;      // during a sweep: upper bounds of every module's T and P (fixed point)
  DevBuf<double> qv;                    // exit probability of every module (0 when empty), during a sweep
  DevBuf<int32_t> mNodes, mUnits, mStale;  // original nodes, units at the current level, stale count
  // one bit per module: it holds exactly one unit (mUnits == 1), rebuilt at
  // the start of every sweep and kept by the commits' apply pass; 4 MB at
  // 33.5M modules, so the singleton rule's reads hit the cache
  DevBuf<uint32_t> mOne;
  DevBuf<uint32_t> mMulti;   // node level: the module held more than one node (mStale > 1) as the sweep started

  // A batch of independent instances (rmx_batch_upload): instance I owns the
  // node ids [b_off[I], b_off[I + 1]) and no arc joins two instances.
  // Modules never leave their instance (an empty module is drawn from the
  // instance's own ids), so every per-instance reduction runs over an id
  // range, and units of every level stay grouped by instance.
  int32_t nb = 0;                          // instances; 0: one network
  std::vector<uint32_t> h_boff;            // node offsets, host copy
  DevBuf<uint32_t> b_off, b_node;          // node offsets (nb + 1); instance of every node
  DevBuf<uint64_t> b_seed;                 // per-instance seed of the batch draw
  DevBuf<double> b_S, b_L, b_C;            // per instance: sum of exits, codelength, sum phi(p)
  DevBuf<double> b_Lold, b_Lstep, b_Louter;  // codelength before a sweep / a super step / an outer round
  DevBuf<int32_t> b_nmod;                  // non-empty modules per instance
  DevBuf<double> b_D, b_R;                 // steady state: dangling mass and L1 residual per instance
  DevBuf<uint32_t> b_iters;                // steady-state iterations per instance
  DevBuf<uint8_t> b_step, b_outer, b_inner;   // still in the super-step / outer / sweep loop
  DevBuf<uint32_t> b_ebeg, b_ecnt, b_ecur; // empty modules per instance: list segment, count, taken
  DevBuf<uint32_t> b_count;                // counters read by the host driver
  // per-instance reductions in chunks of at most kTotChunk ids, none across
  // two instances: chunk k is [b_cbeg[k], b_cend[k]), instance I owns the
  // chunks [b_ichunk[I], b_ichunk[I + 1])
  int32_t n_chunks = 0;
  DevBuf<uint32_t> b_cbeg, b_cend, b_ichunk;
  DevBuf<double> b_cpart;                  // 4 partial sums per chunk

  // sweep work arrays (sized for level 0, reused by level 1)
  DevBuf<uint8_t> act, act_next;
  DevBuf<uint8_t> wmark;             // units with a better module this sweep (neighbour marks by pull)
  DevBuf<uint32_t> target, newmod, stamp, movers, empty_list, want_empty, rank_tmp;
  DevBuf<long long> agg_buf;         // per mover: flows to/from its old and new module
  DevBuf<uint32_t> batch_keys, batch_list;   // this sweep's units by (bucket, batch)
  DevBuf<uint32_t> hub_keys;         // global hash tables of hub units
  DevBuf<long long> hub_out, hub_in;
  DevBuf<uint32_t> hub_plan, hub_bm;  // chunk prefixes of the hubs in a batch; best module per chunk
  DevBuf<double> hub_bd;              // best dL per chunk
  DevBuf<uint32_t> big_list, big_plan;  // commit: movers with hub degree and their arc-chunk prefix
  // partitioned hub evaluation: per batch hub arc, per partition, per hub
  DevBuf<uint32_t> hp_mtmp, hp_ekey;
  DevBuf<long long> hp_ew;
  DevBuf<uint32_t> hp_bcnt, hp_bcur, hp_boff, hp_bm;
  DevBuf<double> hp_bd;
  DevBuf<long long> hp_bfo, hp_bfi;
  DevBuf<uint32_t> hp_plan, hp_ovf, hp_hubu, hp_chub, hp_phub;
  DevBuf<long long> hp_own;
  DevBuf<uint32_t> counters;         // small device counters, then kStripes copies of the launch counters
  DevBuf<uint32_t> movers_st;        // movers per counter copy, mv_cap each
  uint32_t mv_cap = 0;
  DevBuf<double> scal;               // device scalars: S, L, ...
  DevBuf<double> partial;            // block partials for reductions
  uint32_t stamp_now = 0;

  // coarse tune (sub-module partition of every module as its own instance):
  // the intra-module level, sub-module statistics, per-instance totals
  int sub_mode = 0;                  // sweeps run on the sub-module instances
  Level Lsub;
  DevBuf<Stat4> sub_ms;
  DevBuf<int32_t> sub_nodes, sub_units, sub_stale;
  DevBuf<uint32_t> sub_par;          // instance (parent module) of every node
  DevBuf<uint32_t> inst_nodes, inst_start;   // nodes grouped by instance, segment starts
  int32_t n_inst = 0;                // instances (non-empty parent modules)
  DevBuf<double> inst_S, inst_L, inst_Lprev, inst_C;   // per parent module id
  DevBuf<uint8_t> inst_act;
  // instances too large for one wavefront: reduced in chunks (coarse.hip)
  DevBuf<uint32_t> ibig_ctr, ibig_list, ibig_base, ibig_owner;
  DevBuf<double> ibig_part;
  DevBuf<uint32_t> inst_multi;       // the other instances of more than one node
  uint32_t n_ibig = 0, n_ibig_slots = 0, n_multi = 0;
  int32_t n_sub = 0;                 // sub-modules after the last partition

  // steady-state buffers
  DevBuf<double> x, y;
  DevBuf<float> xf;
  DevBuf<double> node_q;   // singleton exits of Module.cpp:86-95 (reporting)

  // tracked scalars (host copies); the initial values are the literal FP64
  // calibrate of Module.cpp:474-501 over the singleton node exits
  double codelength = 0, sum_exit = 0, all_log_all = 0;
  double initial_codelength = 0, initial_sum_exit = 0;
  int iterations = 0;
  double resid = 0;
  int32_t n_module = 0;
  int64_t n_dangling = 0;

  Scratch scratch, scratch2;
  BlockCache cache;   // temporaries of contraction, sorting and codelength
  uint64_t graph_version = 0;   // bumped by every upload
  PushBins pb;                  // the steady state's layout, for graph version pb_graph
  uint64_t pb_graph = 0;
  Timing timing;
  Profiler prof;
  uint64_t sweep_counter = 0;

  // pinned host staging for per-sweep scalars
  double *h_scal = nullptr;
  uint32_t *h_cnt = nullptr;

  Level &cur() { return level == 0 ? L0 : L1; }
  void sync() { RMX_HIP(hipStreamSynchronize(stream)); }
};

// ----- graph.hip
void upload_graph(Ctx &c, int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx,
                  const double *w, const double *node_w);
// A batch of G networks (a union out-CSR, network g owning rows goff[g]..)
// under R seeds each: the networks are uploaded once and copied on the
// device into the union of the G * R instances (instance g * R + r is a copy
// of network g), level 0 then equal to what upload_graph builds from that
// union; L0.uinst holds the instance of every node.
void upload_graph_replicated(Ctx &c, int32_t G, const int32_t *goff, int32_t R, int32_t n, int64_t m,
                             const int64_t *out_ptr, const int32_t *out_idx, const double *w, const double *node_w);
void build_in_csr(Ctx &c, Level &L, const double *raw, DevBuf<double> *raw_in);
void super_in_weights(Ctx &c, const Level &L, double *in_w);   // FP64 in-weights of a contracted level
void build_buckets(Ctx &c, Level &L);

// ----- flow.hip
void steady_state(Ctx &c, int max_iter, int min_iter, double tol);
void init_modules_singletons(Ctx &c);

// ----- codelength.hip
// Recomputes S = sum_m q_m and L from module statistics (tracked values).
void module_totals(Ctx &c, double *S, double *L);
void module_totals_async(Ctx &c);   // leaves S in c.scal[0], L in c.scal[1]
// qv from the module statistics, and S = sum of qv into c.scal[0] (the same
// sum, in the same order, as module_totals_async's S)
void exit_probs_refresh(Ctx &c);
void sum_exit_async(Ctx &c);
void set_partition(Ctx &c, const uint32_t *d_part);
double calculate_codelength(Ctx &c);
void rebuild_exit_flows(Ctx &c);
// Network::calculateCodeLength on raw device arrays (the node level): the
// temporaries of P5Work hold n elements each, `partial` p5_partial_doubles()
// and `scal` 8; the result is left in scal[3] (L, nats), scal[0] (S) and
// scal[7] (non-empty modules).  Sum phi(p) is all_log_all, or, with
// device_log_all, summed from p on the device in the initial calibrate's order.
struct P5Work {
  uint32_t *key, *node, *flag, *pos, *start, *total;
  double *ext, *segq, *segqp, *partial, *scal;
};
size_t p5_partial_doubles();
void p5_launch(int32_t n, const uint32_t *optr, const uint32_t *oidx, const double *ow, const uint32_t *mod,
               const double *p, const double *tau, const uint8_t *dang, double all_log_all, bool device_log_all,
               const P5Work &w, Scratch &scratch, hipStream_t s);

// ----- move.hip
void ensure_sweep_buffers(Ctx &c);
int64_t move_sweep(Ctx &c, bool prioritized, double vthresh, uint64_t seed, int64_t *edge_visits);
void reset_active(Ctx &c);
// One decision pass over every unit of the current level against the same
// module statistics (one batch, nothing committed), as the per-node decision
// of prioritize_parallelMove / parallelMove (Module.cpp:1355-1488): each
// unit's best candidate module (0xffffffff for none; N + 1 for the empty
// module) and its dL in nats, on the device (n units each).
void evaluate_moves(Ctx &c, uint32_t *d_best_m, double *d_best_d);
int64_t n_active(Ctx &c);
// Sub-sweeps of one sweep of level L: the count set on the context, or, when
// it is 0, one per 32768 units of the level, at least 4 and at most 8.  Each
// sub-sweep costs about 20 dependent launches and a host wait (about
// 0.15 ms), which bounds the sweeps of levels up to a few hundred thousand
// units; larger levels keep 8, where fewer sub-sweeps raise the codelength.
// A batch of instances counts the units of its average instance, so that
// attempts run as copies of one network sweep as that network alone does.
// Sub-module partitions keep 8: their instances are single modules of a few
// units, which in fewer sub-sweeps move together against stale statistics
// and can end above their singletons' codelength.
inline int sweep_batches(const Ctx &c, const Level &L) {
  if (c.batches > 0) return c.batches;
  if (c.sub_mode) return 8;
  const int64_t units = c.nb > 0 ? (int64_t)L.n / c.nb : (int64_t)L.n;
  return (int)std::min<int64_t>(8, std::max<int64_t>(4, units >> 15));
}

// ----- coarse.hip
// generate_sub_modules (OmpRelaxmap.cpp:500-632): every module's induced
// subgraph is partitioned as its own network; returns the sub-module count.
int32_t generate_sub_modules(Ctx &c, double threshold, int max_iter, uint64_t seed);
// Network::generateSuperNodesFromSubModules (Module.cpp:3083-3158)
void contract_sub_modules(Ctx &c);
// per-instance S and L of the sub-module partition: of every instance (all),
// else of those with more than one node (the others cannot change)
void instance_totals_async(Ctx &c, bool all = false);

// ----- contract.hip
void update_members(Ctx &c);
void contract(Ctx &c);
// Times every eligible evaluation kernel on the current level's units of
// each degree class and returns the fastest cut-offs that keep buckets in
// kernel order (move.hip).  Evaluation only: the state is left unchanged.
BucketCuts autotune_buckets(Ctx &c);
void propagate_labels(Ctx &c);   // level 1 labels back to nodes
void to_node_level(Ctx &c);      // api.cpp

}  // namespace rmx
