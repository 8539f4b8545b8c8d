/* relaxmap_amd synthetic networks (host, OpenMP): the seeded input graphs of
 * the benchmark and tests.  No counterpart in the reference, which reads
 * networks from files only; rmx_syn_write emits the reference's .net / .txt
 * formats so the reference binary can read the same graphs.
 *
 * Every generator is deterministic for a given seed and thread count
 * independent (one counter-based random stream per node or arc).  Outputs are
 * out-CSR arrays (arcs sorted by (src, dst), no self-loops, duplicates merged
 * by summing weights) ready for rmx_graph_upload, plus planted community
 * labels where the model has them (-1 otherwise). */
#ifndef RELAXMAP_SYNTH_H
#define RELAXMAP_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rmx_syn rmx_syn;

/* Directed stochastic block model: n nodes in `blocks` equal blocks,
 * out-degree ~ Poisson(mean_deg) (distinct targets), each target inside the
 * block with probability 1 - mu, else uniform outside; unit weights. */
int rmx_syn_sbm(int32_t n, int32_t blocks, double mean_deg, double mu, uint64_t seed, rmx_syn **out);

/* LFR-style directed graph: out-degrees power law (exponent deg_exp in
 * [deg_min, deg_max]), community sizes power law (com_exp in [com_min,
 * com_max]), a fraction mu of each node's arcs to uniform targets outside its
 * community, the rest uniform inside; a fraction `dangling` of nodes has no
 * out-arcs; unit weights. */
int rmx_syn_lfr(int32_t n, double deg_exp, int32_t deg_min, int32_t deg_max, double com_exp, int32_t com_min,
                int32_t com_max, double mu, double dangling, uint64_t seed, rmx_syn **out);

/* R-MAT: 2^scale nodes, edge_factor * 2^scale sampled arcs with quadrant
 * probabilities (a, b, c, 1-a-b-c); duplicates summed, self-loops dropped;
 * weights FP32 uniform in (0, 1]; every node kept. */
int rmx_syn_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, rmx_syn **out);

/* Directed Chung-Lu graph with planted communities: out- and in-weights
 * power law (deg_exp in [deg_min, deg_max]) scaled so the mean out-degree is
 * mean_deg (degrees below deg_min are raised to deg_min); targets drawn
 * proportionally to in-weight, inside the community with probability
 * 1 - mu, else from the whole graph outside it; community sizes power law
 * (com_exp in [com_min, com_max]); weights FP32 lognormal(0, 1); a fraction
 * `dangling` of nodes has no out-arcs; node ids randomly permuted. */
int rmx_syn_chunglu(int32_t n, double deg_exp, int32_t deg_min, int32_t deg_max, double mean_deg, double com_exp,
                    int32_t com_min, int32_t com_max, double mu, double dangling, uint64_t seed, rmx_syn **out);

/* Disjoint union of graphs (node ids offset in order); communities offset. */
int rmx_syn_union(rmx_syn *const *parts, int32_t k, rmx_syn **out);

int rmx_syn_info(const rmx_syn *g, int32_t *n, int64_t *m);
/* Borrowed views, valid until rmx_syn_free. */
int rmx_syn_arrays(const rmx_syn *g, const int64_t **out_ptr, const int32_t **out_idx, const double **w,
                   const int32_t **community);
/* Writes .net (Pajek, 1-based, every vertex listed) or .txt (1-based edge
 * list) by the path's extension. */
int rmx_syn_write(const rmx_syn *g, const char *path);
void rmx_syn_free(rmx_syn *g);

#ifdef __cplusplus
}
#endif
#endif
