/* CPU restatement of the reference RelaxMap hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The relaxmap_amd product
 * never links or calls this library.
 *
 * Each function follows one reference function with the same floating-point
 * operation order, so that at one thread (the reference's numTh == 1 path)
 * results are bit-identical to the reference's.  Reference citations are
 * file:line into uwescience/RelaxMap.
 */
#ifndef RELAXMAP_ORACLE_H
#define RELAXMAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_net orc_net;

/* Loaders: FileIO.cpp:32-171 (.net) and :177-265 (.txt), then the network
 * build of OmpRelaxmap.cpp:106-145 (teleport weights, self-loops and w <= 0
 * dropped).  Returns NULL and fills err on the inputs where the reference
 * exits. */
orc_net *orc_load(const char *path, char *err, int errlen);

/* Builds a network from an arc list, as if read from a .net file with n
 * vertices (ids 0..n-1, node weights node_w or 1.0 if NULL).  Duplicate arcs
 * are summed in input order. */
orc_net *orc_from_arcs(int n, int64_t m, const int32_t *src, const int32_t *dst,
                       const double *w, const double *node_w);
void orc_free(orc_net *net);

int orc_n(const orc_net *net);
int64_t orc_m(const orc_net *net);           /* arcs after dedup/self-loop drop */
int orc_n_selflinks(const orc_net *net);
int orc_n_double_links(const orc_net *net);
int orc_name(const orc_net *net, int i, char *buf, int len);

/* Network::initiate + calculateSteadyState + in-link build
 * (Module.cpp:210-469, OmpRelaxmap.cpp:155-162). */
int orc_initiate(orc_net *net);
int orc_iterations(const orc_net *net);
int orc_n_dangling(const orc_net *net);
void orc_get_flows(const orc_net *net, double *p);
void orc_get_node_exit(const orc_net *net, double *q);
void orc_get_teleport(const orc_net *net, double *tau);
/* out-CSR (after initiate: weights are flows f_ab = p_a * w_ab / sum_b w_ab) */
void orc_get_out_csr(const orc_net *net, int64_t *ptr, int32_t *idx, double *w);
void orc_get_raw_out_w(const orc_net *net, double *w);   /* weights as loaded */

double orc_codelength(const orc_net *net);   /* tracked codeLength (nats) */
double orc_sum_exit(const orc_net *net);     /* sumAllExitPr */
double orc_all_log_all(const orc_net *net);
int orc_n_module(const orc_net *net);

/* Installs a partition (module index per node): per-module sums in
 * ascending node order, updateMembersInModule, updateCodeLength
 * (Module.cpp:2881-2962). */
int orc_set_partition(orc_net *net, const int32_t *part);
void orc_get_partition(const orc_net *net, int32_t *part);
void orc_get_module_stats(const orc_net *net, double *exit, double *sum_pr,
                          double *sum_tp, double *sum_dangling, int32_t *n_members);

/* Network::calculateCodeLength (Module.cpp:2966-3002). */
double orc_calculate_codelength(const orc_net *net);
/* Network::updateMembersInModule (Module.cpp:2881-2902). */
void orc_update_members(orc_net *net);

/* Network::convertModulesToSuperNodes (Module.cpp:3008-3077). */
int orc_contract(orc_net *net);
int orc_n_super(const orc_net *net);
int64_t orc_super_m(const orc_net *net);
void orc_get_super(const orc_net *net, double *size, double *exit, double *tp,
                   double *dangling, int32_t *mod, int32_t *n_members);
void orc_get_super_out(const orc_net *net, int64_t *ptr, int32_t *idx, double *w);
void orc_get_super_in(const orc_net *net, int64_t *ptr, int32_t *idx, double *w);

/* MTRand(seed) (MersenneTwister.h:172-183) for the sweep shuffles. */
void orc_seed(orc_net *net, uint32_t seed);
uint32_t orc_rand_int(orc_net *net, uint32_t n);   /* MTRand::randInt(n) */

/* One sweep.  variant: 0 move, 1 prioritize_move, 2 parallelMove (1 thread),
 * 3 prioritize_parallelMove (1 thread).  sp != 0 moves super nodes (the SP
 * twins).  vthresh is the negated CLI vThresh, as the reference passes it.
 * Returns numMoved. */
int orc_sweep(orc_net *net, int variant, int sp, double vthresh);
void orc_reset_active(orc_net *net, int sp);
int orc_n_active(const orc_net *net);

/* generate_network_from_module (OmpRelaxmap.cpp:640-695): the network of
 * the parent's nodes members[0..k) (after orc_initiate; list them ascending,
 * as the reference does) with their flows, teleport weights and dangling
 * flags and the flows of the arcs between them, in singleton modules
 * (Module.cpp:67-99 over those arcs) and calibrated.  partition_module_network
 * (OmpRelaxmap.cpp:436-460) is then orc_sweep(variant 0) until a sweep gains
 * less than the threshold.  NULL on a bad member list. */
orc_net *orc_subnetwork(const orc_net *parent, int k, const int32_t *members);

/* The decision half of prioritize_parallelMove (Module.cpp:1355-1488) for
 * every unit of the level (sp != 0: super nodes) against the same module
 * statistics; nothing moves.  best_mod[u] is the argmin (strict <, ascending
 * module id, the empty-module sentinel n + 1 last) or -1 when no candidate
 * has dL < 0; best_diff[u] its dL (0 for -1).  no_singleton_swaps: a unit
 * alone in its module skips a larger-id module that also holds one unit
 * (not the reference's rule: the GPU's batch rule, for comparison). */
void orc_evaluate_all(orc_net *net, int sp, int no_singleton_swaps, int32_t *best_mod, double *best_diff);
/* Unit u's candidates and their dL as evaluated (its own module left out);
 * returns the count written (at most cap). */
int orc_candidates(orc_net *net, int sp, int u, int cap, int32_t *mods, double *diffs);

/* stochastic_greedy_partition (OmpRelaxmap.cpp:263-429) with numTh == 1
 * and main()'s super-step loop (:181-219) up to the first coarse tune.
 * Returns the number of super steps run; *coarse_next is set when the
 * reference would continue with a coarse tune. */
int orc_run(orc_net *net, double threshold, double vthresh_neg, int max_iter,
            int prior, int *coarse_next);

/* numMoved of each sweep orc_run made (fine and super-node sweeps). */
int orc_trace_len(const orc_net *net);
void orc_get_trace(const orc_net *net, int32_t *moved);
int orc_n_empty(const orc_net *net);

#ifdef __cplusplus
}
#endif
#endif
