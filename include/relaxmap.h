/* relaxmap_amd: C ABI of the MI355X (gfx950) RelaxMap hot path.
 *
 * One rmx_ctx holds one network resident on one GPU, driven by one HIP
 * stream; a context is not thread-safe.  Every call returns 0 on success
 * or a negative status (RMX_E*); rmx_last_error() describes the failure.
 * No C++ exception crosses this boundary.  Arrays are host pointers owned
 * by the caller; ids are 0-based.
 *
 * The reference exposes this path only as member functions of its C++
 * `Network` class (uwescience/RelaxMap Module.h:176-199), called from
 * OmpRelaxmap.cpp.  Each entry point below names the reference function it
 * replaces.
 */
#ifndef RELAXMAP_H
#define RELAXMAP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMX_OK 0
#define RMX_EINVAL (-22)   /* bad argument or malformed input */
#define RMX_ESTATE (-1)    /* call out of order (e.g. sweep before initiate) */
#define RMX_EDEVICE (-19)  /* no usable GPU, or a HIP runtime error */
#define RMX_EIO (-5)       /* file could not be read or written */

#define RMX_FP64 0         /* FP64 throughout */
#define RMX_MIXED 1        /* FP32 arc weights and flows in bulk; exact fixed-point or FP64 sums */

typedef struct rmx_ctx rmx_ctx;

/* Number of visible GPUs (0 if none or no HIP runtime). */
int rmx_device_count(void);

/* Creates a context on `device`.  No reference counterpart (the reference's
 * state is the Network object itself, OmpRelaxmap.cpp:81). */
int rmx_create(rmx_ctx **ctx, int device);
void rmx_destroy(rmx_ctx *ctx);
const char *rmx_last_error(const rmx_ctx *ctx);

/* Precision of the steady state and of the sweeps; call before
 * rmx_initiate (a change after it requires rmx_initiate again: the calls
 * that need flows fail with RMX_ESTATE until then).  Sub-sweeps per sweep:
 * 1..31, or 0 (the default) for one per 32768 units of the level being
 * swept (of its average instance in a batch of instances), at least 4 and
 * at most 8. */
int rmx_set_precision(rmx_ctx *ctx, int precision);
int rmx_set_batches(rmx_ctx *ctx, int batches);

/* How the FP64 steady state iterates (call before rmx_initiate; a change
 * requires rmx_initiate again): 1 (default) the blocked push / pull of the
 * mixed path with FP64 products summed exactly in two-word fixed point,
 * 0 the in-CSR pull with FP64 gathers.  Both stop by the reference's rule
 * and agree with it to its own rounding.  No reference counterpart. */
int rmx_set_fp64_flow(rmx_ctx *ctx, int blocked);

/* How units with more than 2048 arcs are evaluated: 0 (default) partitions
 * their arcs by neighbour module and aggregates each partition in LDS;
 * 1 uses one global-memory hash table per unit (device atomics per arc);
 * 2 partitions, then re-evaluates every hub as 0 does when a partition's
 * table overflows (exercises that fallback).  All give identical moves.
 * No reference counterpart. */
int rmx_set_hub_path(rmx_ctx *ctx, int global_tables);

/* 1: sweeps of a batch of instances (rmx_batch_upload) and of sub-module
 * partitions run without host waits, each sweep's launches captured into one
 * HIP graph and launched at once.  0 (default): the same launches one by one,
 * with a host wait per batch; on MI355X this is the faster of the two (a
 * graph per sweep costs its instantiation and runs no faster).  Both give
 * identical results.  No reference counterpart. */
int rmx_set_graph_capture(rmx_ctx *ctx, int on);

/* The move evaluation's candidate bound for degree bucket `bucket` (0-6 the
 * buckets of rmx_set_bucket_cutoffs, 7 the hub partitions, -1 all): a lower
 * bound of a candidate's dL from the unit's flow to it alone drops the
 * candidates that cannot be a move before their module statistics are read.
 * 1 on, 0 off; the default is on for units of more than 32 arcs (buckets
 * 2-7; those of buckets 0-1 are evaluated unbounded in any case).  Both give
 * identical moves.  No reference counterpart. */
int rmx_set_candidate_bound(rmx_ctx *ctx, int bucket, int mode);

/* How rmx_partition runs attempts > 1 on one network: 0 (default) as one
 * batch of instances (see rmx_batch_upload), a copy of the network per
 * attempt with its own seed, when attempts x arcs <= 2^28 and attempts x nodes
 * <= 2^27, else one after another; 1 always one after another.  Either way
 * every attempt starts from singletons and the partition with the lowest
 * codelength is kept.  No reference counterpart (the reference parses the
 * attempt count and does not use it, OmpRelaxmap.cpp:50). */
int rmx_set_attempt_mode(rmx_ctx *ctx, int sequential);

/* Network construction: replaces the arc-list build of OmpRelaxmap.cpp:113-145
 * (after load_pajek_format_network / load_linkList_format_network,
 * FileIO.cpp:32-265).  out_ptr[n+1] / out_idx[m] / w[m] is the out-CSR with
 * arcs sorted by (src, dst), no self-loops, no duplicates, w > 0; node_w[n]
 * are node (teleport) weights, NULL for 1.0.  m < 2^31 (larger m is refused). */
int rmx_graph_upload(rmx_ctx *ctx, int32_t n, int64_t m, const int64_t *out_ptr, const int32_t *out_idx,
                     const double *w, const double *node_w);

/* Reads a .net (Pajek) or edge-list .txt file with the reference's parsing
 * rules (FileIO.cpp:32-265) and uploads it (rmx_graph_upload).  Node names
 * are kept on the host for the output writers. */
int rmx_load_network(rmx_ctx *ctx, const char *path, int32_t *n, int64_t *m, int32_t *n_self_links,
                     int32_t *n_double_links);

/* Degree buckets of the sweeps (no reference counterpart): units with
 * degree (out + in) in (max_deg[k-1], max_deg[k]] are evaluated by kernel k
 * (k = 0..6: 8..64-lane groups with 32..1024-key LDS tables, then 2048- and
 * 4096-key workgroup tables); above max_deg[6], by the hub partitions.
 * max_deg[k] may not exceed its default (16, 32, 128, 256, 512, 1023,
 * 2048), the largest degree the kernel's table holds at load 1/2.  The
 * choice changes speed only: every kernel finds the same best move.
 * rmx_autotune times the eligible kernels on the current level's units of
 * each degree class (evaluation only, nothing moves) and installs the
 * fastest cut-offs that keep the kernel order. */
#define RMX_BUCKET_CUTS 7
int rmx_set_bucket_cutoffs(rmx_ctx *ctx, const int32_t *max_deg);
int rmx_get_bucket_cutoffs(rmx_ctx *ctx, int32_t *max_deg);
int rmx_autotune(rmx_ctx *ctx, int32_t *max_deg);

/* A batch of independent networks (many graphs, or one graph under many
 * seeds) partitioned together: the union of the networks' out-CSRs, where
 * instance I owns the node ids [node_off[I], node_off[I+1]) and no arc joins
 * two instances.  seeds[I] drives the sweeps of instance I.  After
 * rmx_initiate (steady state of every instance on its own, with the stop
 * rule applied per instance) rmx_partition runs main()'s schedule
 * (OmpRelaxmap.cpp:181-219) for every instance at once, each with its own
 * codelength and stop decisions; modules never leave their instance.  No
 * reference counterpart: the reference runs one network per process.
 * attempts > 1 and rmx_calculate_codelength are refused on a batch. */
int rmx_batch_upload(rmx_ctx *ctx, int32_t n_inst, const int32_t *node_off, const uint64_t *seeds, int32_t n,
                     int64_t m, const int64_t *out_ptr, const int32_t *out_idx, const double *w,
                     const double *node_w);
/* The same batch when every network runs under several seeds (the '#
 * attempts' of OmpRelaxmap.cpp:50 as independent instances): n_graph networks
 * as one union out-CSR (network g owns the node ids [graph_off[g],
 * graph_off[g+1]) and no arc joins two networks), each run under n_seed
 * seeds.  Instance k = g * n_seed + r is network g with seed seeds[k]; its
 * node ids follow those of instance k - 1, as rmx_batch_upload would number
 * the union of the copies.  The networks cross to the device once and are
 * copied there; every later result equals rmx_batch_upload's on that union.
 * No reference counterpart. */
int rmx_batch_upload_seeds(rmx_ctx *ctx, int32_t n_graph, const int32_t *graph_off, int32_t n_seed,
                           const uint64_t *seeds, int32_t n, int64_t m, const int64_t *out_ptr,
                           const int32_t *out_idx, const double *w, const double *node_w);
/* Per instance: codelength (nats) and non-empty modules of the current
 * partition, and the steady-state iterations. */
int rmx_batch_results(rmx_ctx *ctx, double *codelength, int32_t *n_module, int32_t *iterations);

/* Host-side network files, no GPU needed.  rmx_net_read parses a .net or
 * .txt file with the reference's rules (FileIO.cpp:32-265, then the
 * self-loop and w <= 0 removal of OmpRelaxmap.cpp:117-135) on `threads` host
 * threads (0: all), or reads a .rmxcsr binary written by rmx_net_write.  The
 * arrays stay valid until rmx_net_free.  No reference counterpart. */
typedef struct rmx_net rmx_net;
int rmx_net_read(const char *path, int threads, rmx_net **net, char *err, int err_len);
void rmx_net_free(rmx_net *net);
int rmx_net_info(const rmx_net *net, int32_t *n, int64_t *m, int32_t *n_self_links, int32_t *n_double_links);
int rmx_net_arrays(const rmx_net *net, const int64_t **out_ptr, const int32_t **out_idx, const double **w,
                   const double **node_w);
const char *rmx_net_name(const rmx_net *net, int32_t i);
int rmx_net_write(const rmx_net *net, const char *path);
/* Uploads a parsed network (rmx_graph_upload) and keeps its names for the writers. */
int rmx_net_upload(rmx_ctx *ctx, const rmx_net *net);

/* Network::initiate (Module.cpp:210-333): dangling nodes, row
 * normalisation, steady state (Network::calculateSteadyState,
 * Module.cpp:340-469: at least 50, at most 200 iterations, L1 change
 * <= 1e-15), arc flows, singleton modules, calibrate (Module.cpp:474-501).
 * Outputs may be NULL.  *codelength is in nats. */
int rmx_initiate(rmx_ctx *ctx, int32_t *iterations, double *codelength, int64_t *n_dangling);
/* Steady-state node flows p (Node::Size after initiate). */
int rmx_get_flows(rmx_ctx *ctx, double *p);
/* Singleton exit probabilities (Module::Module(int, Node*), Module.cpp:86-95). */
int rmx_get_node_exit(rmx_ctx *ctx, double *q);

/* Tracked codelength, sum of module exits and module count of the current
 * partition (Network::CodeLength / SumAllExitPr / NModule). */
int rmx_codelength(rmx_ctx *ctx, double *codelength, double *sum_exit, int32_t *n_module);
/* Network::calculateCodeLength (Module.cpp:2966-3002): from scratch, FP64. */
int rmx_calculate_codelength(rmx_ctx *ctx, double *codelength);

/* Installs a node partition (module index per node, 0..n-1) with freshly
 * summed module statistics (as Network::updateMembersInModule +
 * updateCodeLength, Module.cpp:2881-2962, do for a given assignment). */
int rmx_set_partition(rmx_ctx *ctx, const int32_t *part);
/* Module index per node (Node::ModIdx). */
int rmx_get_partition(rmx_ctx *ctx, int32_t *part);

/* All units active again (OmpRelaxmap.cpp:282-288, 370-376). */
int rmx_reset_active(rmx_ctx *ctx);
/* One sweep over the current level's units: Network::prioritize_parallelMove
 * / parallelMove (Module.cpp:962-1685) on nodes, and their super-node twins
 * (Module.cpp:2154-2869) after rmx_contract.  vthresh is the CLI vThresh
 * (moves need dL < -vthresh when prioritized, dL < 0 otherwise).  *moved
 * counts original nodes moved.  Outputs may be NULL. */
int rmx_move_sweep(rmx_ctx *ctx, int prioritized, double vthresh, uint64_t seed, int64_t *moved,
                   double *codelength, int64_t *edge_visits);

/* The decision half of Network::prioritize_parallelMove / parallelMove
 * (Module.cpp:1355-1488, and the super-node twins Module.cpp:2154-2869
 * after rmx_contract): every unit of the current level is evaluated against
 * the same module statistics, as one batch, and nothing is committed.
 * best_module[u] is the candidate the unit would move to (strict < over the
 * neighbour modules in ascending id, the empty-module option, offered as in
 * Module.cpp:1435-1439, as id N + 1 after them), or -1 when no candidate
 * lowers the codelength; best_delta[u] is its delta codelength in nats (0
 * for -1).  The sweeps' own batch rule applies: a unit alone in its module
 * does not take another such module of larger id.  Not on a batch of
 * instances.  The arrays hold `capacity` entries (N always suffices);
 * *n_units receives the level's unit count. */
int rmx_evaluate_moves(rmx_ctx *ctx, int32_t capacity, int32_t *n_units, int32_t *best_module,
                       double *best_delta);

/* Network::updateMembersInModule (Module.cpp:2881-2902). */
int rmx_update_members(rmx_ctx *ctx);
/* Network::convertModulesToSuperNodes (Module.cpp:3008-3077): contracts the
 * current level by the current modules (at the node level the original arcs,
 * as the reference does; after a contraction its super arcs, which gives the
 * same exact fixed-point sums and FP64 weights equal up to rounding); later
 * sweeps move the super nodes. */
int rmx_contract(rmx_ctx *ctx, int32_t *n_super, int64_t *m_super);
/* Super-node statistics after rmx_contract (SuperNode(Module, int),
 * Node.cpp:76-106): flow, exit, teleport, dangling flow, module index and
 * member count.  Any pointer may be NULL. */
int rmx_get_super(rmx_ctx *ctx, double *size, double *exit, double *tp, double *dangling, int32_t *mod,
                  int32_t *n_members);
/* Super-node out- and in-lists (SuperNode::outLinks / inLinks). */
int rmx_get_super_arcs(rmx_ctx *ctx, int64_t *out_ptr, int32_t *out_idx, double *out_w, int64_t *in_ptr,
                       int32_t *in_idx, double *in_w);
/* Back to moving original nodes (node modules follow their super nodes). */
int rmx_to_node_level(rmx_ctx *ctx);

/* generate_sub_modules (OmpRelaxmap.cpp:500-632): every module's induced
 * subgraph (generate_network_from_module, OmpRelaxmap.cpp:640-783) is
 * partitioned from singletons as its own network, with its own codelength
 * (partition_module_network, OmpRelaxmap.cpp:436-498: sweeps until one gains
 * less than `threshold` bits, at most max_iter).  All modules are solved
 * together on the GPU.  *n_sub is the number of sub-modules. */
int rmx_generate_sub_modules(rmx_ctx *ctx, double threshold, int32_t max_iter, uint64_t seed, int32_t *n_sub);
/* Sub-module label per node after rmx_generate_sub_modules (a node id of the
 * same parent module). */
int rmx_get_sub_modules(rmx_ctx *ctx, int32_t *sub);
/* Network::generateSuperNodesFromSubModules (Module.cpp:3083-3158): every
 * sub-module becomes a super node in its parent module; later sweeps move
 * them (the coarse tune).  The super-node exit is the consistent
 * (alpha P + beta D)(1 - T) + beta X, not the reference's per-member sum. */
int rmx_contract_sub_modules(rmx_ctx *ctx, int32_t *n_super, int64_t *m_super);

/* The CLI's optimisation: main()'s super-step loop (OmpRelaxmap.cpp:181-219)
 * around stochastic_greedy_partition (OmpRelaxmap.cpp:263-429). */
typedef struct {
  double threshold;   /* stop when a super step gains less (bits) */
  double vthresh;     /* CLI vThresh */
  int32_t max_iter;   /* sweeps per phase */
  int32_t prior;      /* 1: prioritized sweeps ("prior"), 0: "normal" */
  uint64_t seed;
  int32_t verbose;    /* 1: print the reference's progress lines to stdout */
  int32_t attempts;   /* > 1: independent runs from singletons with seeds seed, seed + 1, ...;
                         the partition with the lowest codelength is kept (the CLI's
                         <# attempts>, parsed but unused by the reference, OmpRelaxmap.cpp:50) */
} rmx_partition_opts;

typedef struct {
  double codelength;          /* tracked, nats */
  double computed_codelength; /* Network::calculateCodeLength, nats */
  int32_t n_module;
  int32_t super_steps;
  int64_t sweeps;
  int64_t edge_visits;
  double time_partition_s;    /* sweeps + contraction (OmpRelaxmap.cpp:222) */
  double time_sweeps_s;
  double time_contract_s;
  double time_coarse_s;       /* sub-module partitions of the coarse tunes */
  int32_t best_attempt;       /* the attempt kept (0-based) */
} rmx_partition_result;

int rmx_partition(rmx_ctx *ctx, const rmx_partition_opts *opts, rmx_partition_result *res);

/* Output writers: the .tree (OmpRelaxmap.cpp:790-812) and .clu
 * (OmpRelaxmap.cpp:240-249) files of the current partition.  Node names
 * come from rmx_load_network (decimal 1-based ids after rmx_graph_upload). */
int rmx_write_tree(rmx_ctx *ctx, const char *path);
int rmx_write_clu(rmx_ctx *ctx, const char *path);

/* Profiling mode (off by default): every hot kernel launch is bracketed by
 * HIP events on the context's stream and accounted with its algorithmic HBM
 * bytes.  rmx_kernel_stats(ctx, i, ...) returns entry i (name, launches,
 * total milliseconds, total bytes); with i < 0 it returns the entry count.
 * No reference counterpart (the reference prints wall-clock phases only). */
int rmx_set_profiling(rmx_ctx *ctx, int on);
int rmx_kernel_stats(rmx_ctx *ctx, int i, char *name, int name_len, int64_t *launches, double *ms, double *bytes);

/* No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It computes what rmx_calculate_codelength
 * does (Network::calculateCodeLength, Module.cpp:2966-3002) for a network and
 * partition held in the caller's device memory, on the caller's stream
 * (hipStream_t, NULL for the default stream) of GPU `device`:
 *   d_out_ptr[n+1], d_out_idx[m]  out-CSR (device), heads in [0, n)
 *   d_flow[m]                     arc flows f_ab = p_a * w_ab / sum_b w_ab
 *   d_p[n], d_tau[n]              steady-state flows and teleport weights
 *   d_dangling[n]                 1 for a node without out-arcs, else 0
 *   d_module[n]                   module index of every node, in [0, n)
 * *codelength (host) receives the codelength in nats.  Every temporary lives
 * in a static device array of the library (one per GPU, held until the
 * process ends), sized for n <= RMX_NOALLOC_MAX_NODES; larger n is refused
 * with RMX_EINVAL.  The call allocates nothing, checks the CSR and the module
 * indices on the device first (RMX_EINVAL on a violation), and returns after
 * its stream has finished.  Calls that overlap on one GPU share that array
 * and race.  On failure a message goes to err[err_len] (may be NULL). */
#define RMX_NOALLOC_MAX_NODES (1 << 20)
int rmx_calculate_codelength_noalloc(int device, void *stream, int32_t n, int64_t m, const uint32_t *d_out_ptr,
                                     const uint32_t *d_out_idx, const double *d_flow, const double *d_p,
                                     const double *d_tau, const uint8_t *d_dangling, const int32_t *d_module,
                                     double *codelength, char *err, int err_len);

/* Kernel timings of the last rmx_initiate / rmx_partition (milliseconds). */
int rmx_timing(rmx_ctx *ctx, double *flow_ms, double *sweep_ms, double *contract_ms, int64_t *edge_visits,
               int64_t *sweeps);

#ifdef __cplusplus
}
#endif
#endif
