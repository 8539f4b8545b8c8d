/* CPU restatement of the reference RelaxMap hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library as the checker.  The relaxmap_amd
 * product never links or calls it.
 *
 * Every function restates one reference function (file:line citations into
 * uwescience/RelaxMap) keeping its floating-point operation order, so that
 * at one thread the results are bit-identical to the reference's.  The
 * restatement is plain C: per-node std::map aggregation becomes a small
 * open-addressing table whose keys are visited in ascending order, the
 * Network/Module/SuperNode objects become flat arrays.
 */
#define _POSIX_C_SOURCE 200809L
#include "relaxmap_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Module.h:114-116: beta is the double nearest 1.0 - 0.15, not 0.85. */
static const double ALPHA = 0.15;
static const double BETA = 1.0 - 0.15;

/* Module.h:23-25 */
static double plogp(double x) { return x > 0.0 ? x * log(x) : 0.0; }

/* ------------------------------------------------------------------ */
/* Mersenne Twister, as seeded and drawn by MersenneTwister.h:142-183,  */
/* 233-245 (MT19937 of Matsumoto & Nishimura with the Knuth 69069       */
/* initialisation and bit-mask rejection for bounded integers).         */
/* ------------------------------------------------------------------ */
enum { MT_N = 624, MT_M = 397 };
typedef struct { uint32_t s[MT_N]; int next, left; } mt_state;

static uint32_t mt_twist(uint32_t m, uint32_t s0, uint32_t s1) {
  uint32_t y = (s0 & 0x80000000u) | (s1 & 0x7fffffffu);
  return m ^ (y >> 1) ^ ((s1 & 1u) ? 0x9908b0dfu : 0u);
}

static void mt_reload(mt_state *mt) {
  uint32_t *p = mt->s;
  int i;
  for (i = 0; i < MT_N - MT_M; i++, p++) *p = mt_twist(p[MT_M], p[0], p[1]);
  for (i = 0; i < MT_M - 1; i++, p++) *p = mt_twist(p[MT_M - MT_N], p[0], p[1]);
  *p = mt_twist(p[MT_M - MT_N], p[0], mt->s[0]);
  mt->left = MT_N;
  mt->next = 0;
}

static void mt_seed(mt_state *mt, uint32_t seed) {
  /* The high half of each state word is taken before, the low half after
   * one 69069 step; the "+1" of that step is applied after the low half is
   * read (MersenneTwister.h:177-181). */
  for (int i = 0; i < MT_N; i++) {
    uint32_t v = seed & 0xffff0000u;
    seed *= 69069u;
    v |= (seed & 0xffff0000u) >> 16;
    seed += 1u;
    seed = seed * 69069u + 1u;
    mt->s[i] = v;
  }
  mt_reload(mt);
}

static uint32_t mt_rand(mt_state *mt) {
  if (mt->left == 0) mt_reload(mt);
  --mt->left;
  uint32_t y = mt->s[mt->next++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  return y ^ (y >> 18);
}

static uint32_t mt_rand_n(mt_state *mt, uint32_t n) {
  uint32_t used = n;
  used |= used >> 1; used |= used >> 2; used |= used >> 4;
  used |= used >> 8; used |= used >> 16;
  uint32_t i;
  do { i = mt_rand(mt) & used; } while (i > n);
  return i;
}

/* ------------------------------------------------------------------ */
/* Data                                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  int n;
  int64_t m;
  int64_t *out_ptr, *in_ptr;
  int32_t *out_idx, *in_idx;
  double *out_w, *in_w;
  double *size, *tp, *dsize, *exitpr;
  char *dangling;
  int32_t *mod;
  /* super nodes only: member node lists (CSR) */
  int64_t *mem_ptr;
  int32_t *mem_idx;
} level_t;

struct orc_net {
  int n;
  int64_t m;
  int n_self, n_double;
  double tot_node_w;
  char **names;
  double *node_w;
  double *raw_w;     /* out weights as loaded */
  level_t L;         /* original nodes */
  level_t S;         /* super nodes (after orc_contract) */
  int has_super;
  int iterations, n_dangling;
  int32_t *danglings;
  /* modules (indexed 0..n-1) */
  double *m_exit, *m_stay, *m_sum, *m_tp, *m_dang;
  int32_t *m_n;
  int64_t *mem_ptr;   /* updateMembersInModule lists, ascending node id */
  int32_t *mem_idx;
  int32_t *mem_cnt;   /* members.size() (stale between updates) */
  int32_t *empty;     /* emptyModules stack */
  int n_empty_stack;
  int n_module, n_empty_mod;
  double codelength, sum_exit, all_log_all;
  mt_state mt;
  /* prioritisation */
  char *is_active;
  int32_t *active;
  int n_active;
  /* numMoved of every sweep run by orc_run, in order */
  int32_t *trace;
  int n_trace, trace_cap;
  /* orc_evaluate_all / orc_candidates: units per module for the batch rule
   * (NULL: not applied), and the candidate record (NULL: not recorded) */
  int32_t *eval_units;
  int32_t *cand_mod;
  double *cand_diff;
  int cand_n, cand_cap;
  /* per-node scratch for the neighbour-module table */
  int tab_cap;
  int32_t *tab_key;
  double *tab_out, *tab_in;
  int32_t *tab_used, *tab_order;
};

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

static void level_free(level_t *L) {
  free(L->out_ptr); free(L->in_ptr); free(L->out_idx); free(L->in_idx);
  free(L->out_w); free(L->in_w); free(L->size); free(L->tp); free(L->dsize);
  free(L->exitpr); free(L->dangling); free(L->mod); free(L->mem_ptr); free(L->mem_idx);
  memset(L, 0, sizeof(*L));
}

void orc_free(orc_net *net) {
  if (!net) return;
  if (net->names) { for (int i = 0; i < net->n; i++) free(net->names[i]); free(net->names); }
  free(net->node_w); free(net->raw_w);
  level_free(&net->L); level_free(&net->S);
  free(net->danglings);
  free(net->m_exit); free(net->m_stay); free(net->m_sum); free(net->m_tp); free(net->m_dang);
  free(net->m_n); free(net->mem_ptr); free(net->mem_idx); free(net->mem_cnt); free(net->empty);
  free(net->is_active); free(net->active); free(net->trace);
  free(net->tab_key); free(net->tab_out); free(net->tab_in); free(net->tab_used); free(net->tab_order);
  free(net);
}

/* ------------------------------------------------------------------ */
/* Arc list -> network (OmpRelaxmap.cpp:113-145)                        */
/* ------------------------------------------------------------------ */
typedef struct { int64_t a, b; double w; int64_t seq; } arc_t;

static int arc_cmp(const void *x, const void *y) {
  const arc_t *p = (const arc_t *)x, *q = (const arc_t *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  return p->seq < q->seq ? -1 : (p->seq > q->seq);
}

/* Sorts arcs by (a, b) and sums duplicates in input order, as
 * `Edges[make_pair(a, b)] += w` does (FileIO.cpp:154, :220).  Returns the
 * number of unique arcs; *n_double counts additions that found a positive
 * running sum. */
static int64_t merge_arcs(arc_t *arcs, int64_t m, int *n_double) {
  qsort(arcs, (size_t)m, sizeof(arc_t), arc_cmp);
  int64_t u = 0;
  *n_double = 0;
  for (int64_t i = 0; i < m; i++) {
    if (u > 0 && arcs[u - 1].a == arcs[i].a && arcs[u - 1].b == arcs[i].b) {
      double s = arcs[u - 1].w + arcs[i].w;
      if (s > arcs[i].w) (*n_double)++;
      arcs[u - 1].w = s;
    } else {
      double s = 0.0 + arcs[i].w;
      if (s > arcs[i].w) (*n_double)++;
      arcs[u] = arcs[i];
      arcs[u].w = s;
      u++;
    }
  }
  return u;
}

static orc_net *net_alloc(int n) {
  orc_net *net = (orc_net *)xcalloc(1, sizeof(orc_net));
  net->n = n;
  net->names = (char **)xcalloc(n, sizeof(char *));
  net->node_w = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; i++) net->node_w[i] = 1.0;
  return net;
}

/* Builds out-links from unique sorted arcs in 0..n-1 ids: w > 0 kept,
 * self-loops (w > 0) counted and dropped (OmpRelaxmap.cpp:117-135). */
static void build_out(orc_net *net, const arc_t *arcs, int64_t u) {
  int n = net->n;
  level_t *L = &net->L;
  L->n = n;
  L->out_ptr = (int64_t *)xcalloc(n + 1, sizeof(int64_t));
  int64_t m = 0;
  net->n_self = 0;
  for (int64_t i = 0; i < u; i++) {
    if (arcs[i].w > 0.0) {
      if (arcs[i].a == arcs[i].b) net->n_self++;
      else { L->out_ptr[arcs[i].a + 1]++; m++; }
    }
  }
  for (int i = 0; i < n; i++) L->out_ptr[i + 1] += L->out_ptr[i];
  L->out_idx = (int32_t *)xcalloc(m, sizeof(int32_t));
  L->out_w = (double *)xcalloc(m, sizeof(double));
  net->raw_w = (double *)xcalloc(m, sizeof(double));
  int64_t k = 0;
  for (int64_t i = 0; i < u; i++) {
    if (arcs[i].w > 0.0 && arcs[i].a != arcs[i].b) {
      L->out_idx[k] = (int32_t)arcs[i].b;
      L->out_w[k] = arcs[i].w;
      net->raw_w[k] = arcs[i].w;
      k++;
    }
  }
  L->m = m;
  net->m = m;
  double tot = 0.0;
  for (int i = 0; i < n; i++) tot += net->node_w[i];
  net->tot_node_w = tot;
  L->tp = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; i++) L->tp[i] = net->node_w[i] / tot;
}

orc_net *orc_from_arcs(int n, int64_t m, const int32_t *src, const int32_t *dst,
                       const double *w, const double *node_w) {
  if (n <= 0) return NULL;
  arc_t *arcs = (arc_t *)xcalloc(m, sizeof(arc_t));
  for (int64_t i = 0; i < m; i++) {
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) { free(arcs); return NULL; }
    arcs[i].a = src[i]; arcs[i].b = dst[i]; arcs[i].w = w ? w[i] : 1.0; arcs[i].seq = i;
  }
  orc_net *net = net_alloc(n);
  char buf[32];
  for (int i = 0; i < n; i++) {
    snprintf(buf, sizeof buf, "%d", i + 1);
    net->names[i] = strdup(buf);
    if (node_w) net->node_w[i] = node_w[i] > 0.0 ? node_w[i] : 1.0;
  }
  int64_t u = merge_arcs(arcs, m, &net->n_double);
  build_out(net, arcs, u);
  free(arcs);
  return net;
}

/* ------------------------------------------------------------------ */
/* Loaders (FileIO.cpp:32-265).  Tokens follow `istringstream >> string` */
/* semantics: whitespace-separated, and a failed extraction leaves the  */
/* previous token in place (FileIO.cpp:143-153, 202-214).               */
/* ------------------------------------------------------------------ */
typedef struct { const char *p, *end; int failed; } tok_t;

static void tok_init(tok_t *t, const char *p, const char *end) { t->p = p; t->end = end; t->failed = 0; }

/* Reads the next token into buf (unchanged on failure). */
static int tok_next(tok_t *t, char *buf, size_t cap) {
  if (t->failed) return 0;
  while (t->p < t->end && isspace((unsigned char)*t->p)) t->p++;
  if (t->p >= t->end) { t->failed = 1; return 0; }
  size_t k = 0;
  while (t->p < t->end && !isspace((unsigned char)*t->p)) {
    if (k + 1 < cap) buf[k++] = *t->p;
    t->p++;
  }
  buf[k] = 0;
  return 1;
}

static char *read_file(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *d = (char *)malloc((size_t)sz + 1);
  if (!d) { fclose(f); return NULL; }
  size_t got = fread(d, 1, (size_t)sz, f);
  fclose(f);
  d[got] = 0;
  *len = got;
  return d;
}

/* std::getline over '\n'; the last line needs no terminator. */
typedef struct { const char *p, *end; } lines_t;
static int next_line(lines_t *ls, const char **b, const char **e) {
  if (ls->p >= ls->end) return 0;
  *b = ls->p;
  const char *q = memchr(ls->p, '\n', (size_t)(ls->end - ls->p));
  if (q) { *e = q; ls->p = q + 1; } else { *e = ls->end; ls->p = ls->end; }
  return 1;
}

static int64_t cmp_i64(int64_t a, int64_t b) { return (a > b) - (a < b); }
static int i64_cmp(const void *x, const void *y) { return (int)cmp_i64(*(const int64_t *)x, *(const int64_t *)y); }

static orc_net *load_txt(const char *data, size_t len, char *err, int errlen) {
  lines_t ls = {data, data + len};
  const char *b, *e;
  char buf[256] = "";
  int64_t cap = 1024, m = 0;
  arc_t *arcs = (arc_t *)xcalloc(cap, sizeof(arc_t));
  while (next_line(&ls, &b, &e)) {
    if (b < e && *b == '#') continue;
    tok_t t;
    tok_init(&t, b, e);
    tok_next(&t, buf, sizeof buf);
    int64_t a = atol(buf);
    tok_next(&t, buf, sizeof buf);
    int64_t c = atol(buf);
    buf[0] = 0;
    tok_next(&t, buf, sizeof buf);
    double w = buf[0] ? atof(buf) : 1.0;
    if (m == cap) { cap *= 2; arcs = (arc_t *)realloc(arcs, (size_t)cap * sizeof(arc_t)); }
    arcs[m].a = a; arcs[m].b = c; arcs[m].w = w; arcs[m].seq = m;
    m++;
  }
  /* node set = union of endpoints, renumbered in ascending id order */
  int64_t *ids = (int64_t *)xcalloc(2 * m, sizeof(int64_t));
  for (int64_t i = 0; i < m; i++) { ids[2 * i] = arcs[i].a; ids[2 * i + 1] = arcs[i].b; }
  qsort(ids, (size_t)(2 * m), sizeof(int64_t), i64_cmp);
  int64_t nu = 0;
  for (int64_t i = 0; i < 2 * m; i++) if (nu == 0 || ids[nu - 1] != ids[i]) ids[nu++] = ids[i];
  if (nu == 0) {
    snprintf(err, errlen, "network has no nodes");
    free(ids); free(arcs);
    return NULL;
  }
  orc_net *net = net_alloc((int)nu);
  char nb[32];
  for (int64_t i = 0; i < nu; i++) { snprintf(nb, sizeof nb, "%ld", (long)ids[i]); net->names[i] = strdup(nb); }
  for (int64_t i = 0; i < m; i++) {
    int64_t *pa = bsearch(&arcs[i].a, ids, (size_t)nu, sizeof(int64_t), i64_cmp);
    int64_t *pb = bsearch(&arcs[i].b, ids, (size_t)nu, sizeof(int64_t), i64_cmp);
    arcs[i].a = pa - ids; arcs[i].b = pb - ids;
  }
  free(ids);
  int64_t u = merge_arcs(arcs, m, &net->n_double);
  build_out(net, arcs, u);
  free(arcs);
  return net;
}

static int is_vertices(const char *s) { return !strcmp(s, "*Vertices") || !strcmp(s, "*vertices") || !strcmp(s, "*VERTICES"); }
static int is_arcs(const char *s) { return !strcmp(s, "*Edges") || !strcmp(s, "*edges") || !strcmp(s, "*Arcs") || !strcmp(s, "*arcs"); }

static orc_net *load_net(const char *data, size_t len, char *err, int errlen) {
  lines_t ls = {data, data + len};
  const char *b, *e;
  char buf[256] = "";
  int n = 0;
  while (n == 0) {
    if (!next_line(&ls, &b, &e)) { snprintf(err, errlen, "the network file is not in Pajek format"); return NULL; }
    tok_t t;
    tok_init(&t, b, e);
    tok_next(&t, buf, sizeof buf);
    if (is_vertices(buf)) { tok_next(&t, buf, sizeof buf); n = atoi(buf); }
    else { snprintf(err, errlen, "the network file is not in Pajek format"); return NULL; }
  }
  if (n < 0) { snprintf(err, errlen, "negative vertex count"); return NULL; }
  orc_net *net = net_alloc(n);
  for (int i = 0; i < n; i++) {
    if (!next_line(&ls, &b, &e)) { b = e = data + len; }
    tok_t t;
    tok_init(&t, b, e);
    tok_next(&t, buf, sizeof buf);   /* id, ignored (FileIO.cpp:72-75) */
    const char *q1 = NULL, *q2 = NULL;
    for (const char *c = b; c < e; c++) if (*c == '"') { if (!q1) q1 = c; q2 = c; }
    if (q1 && q1 < q2) {
      size_t k = (size_t)(q2 - q1 - 1);
      net->names[i] = (char *)xcalloc(k + 1, 1);
      memcpy(net->names[i], q1 + 1, k);
      tok_init(&t, q2 + 1, e);
    } else {
      tok_init(&t, b, e);
      tok_next(&t, buf, sizeof buf);
      tok_next(&t, buf, sizeof buf);
      net->names[i] = strdup(buf);
    }
    strcpy(buf, "1");
    tok_next(&t, buf, sizeof buf);
    double nw = atof(buf);
    if (nw <= 0.0) nw = 1.0;
    net->node_w[i] = nw;
  }
  {
    if (!next_line(&ls, &b, &e)) b = e = data + len;
    tok_t t;
    tok_init(&t, b, e);
    tok_next(&t, buf, sizeof buf);
    if (!is_arcs(buf)) { snprintf(err, errlen, "Number of nodes not matching"); orc_free(net); return NULL; }
  }
  int64_t cap = 1024, m = 0;
  arc_t *arcs = (arc_t *)xcalloc(cap, sizeof(arc_t));
  while (next_line(&ls, &b, &e)) {
    tok_t t;
    tok_init(&t, b, e);
    tok_next(&t, buf, sizeof buf);
    int a = atoi(buf);
    tok_next(&t, buf, sizeof buf);
    int c = atoi(buf);
    buf[0] = 0;
    tok_next(&t, buf, sizeof buf);
    double w = buf[0] ? atof(buf) : 1.0;
    a--; c--;
    if (m == cap) { cap *= 2; arcs = (arc_t *)realloc(arcs, (size_t)cap * sizeof(arc_t)); }
    arcs[m].a = a; arcs[m].b = c; arcs[m].w = w; arcs[m].seq = m;
    m++;
  }
  int64_t u = merge_arcs(arcs, m, &net->n_double);
  /* The reference indexes nodes[] with these ids unchecked (undefined
   * behaviour out of range); self-loops among them are harmless. */
  for (int64_t i = 0; i < u; i++) {
    if (arcs[i].w > 0.0 && arcs[i].a != arcs[i].b &&
        (arcs[i].a < 0 || arcs[i].a >= n || arcs[i].b < 0 || arcs[i].b >= n)) {
      snprintf(err, errlen, "arc %ld -> %ld outside 1..%d", (long)arcs[i].a + 1, (long)arcs[i].b + 1, n);
      free(arcs); orc_free(net);
      return NULL;
    }
  }
  build_out(net, arcs, u);
  free(arcs);
  return net;
}

orc_net *orc_load(const char *path, char *err, int errlen) {
  size_t len = 0;
  char *d = read_file(path, &len);
  if (!d) { snprintf(err, errlen, "cannot open %s", path); return NULL; }
  const char *dot = strrchr(path, '.');
  orc_net *net = (dot && !strcmp(dot, ".net")) ? load_net(d, len, err, errlen) : load_txt(d, len, err, errlen);
  free(d);
  return net;
}

int orc_n(const orc_net *net) { return net->n; }
int64_t orc_m(const orc_net *net) { return net->m; }
int orc_n_selflinks(const orc_net *net) { return net->n_self; }
int orc_n_double_links(const orc_net *net) { return net->n_double; }
int orc_name(const orc_net *net, int i, char *buf, int len) {
  if (i < 0 || i >= net->n) return -1;
  snprintf(buf, len, "%s", net->names[i] ? net->names[i] : "");
  return 0;
}

/* ------------------------------------------------------------------ */
/* Steady state and initial modules (Module.cpp:210-501)                */
/* ------------------------------------------------------------------ */
static void build_in(level_t *L) {
  int n = L->n;
  L->in_ptr = (int64_t *)xcalloc(n + 1, sizeof(int64_t));
  for (int64_t k = 0; k < L->m; k++) L->in_ptr[L->out_idx[k] + 1]++;
  for (int i = 0; i < n; i++) L->in_ptr[i + 1] += L->in_ptr[i];
  L->in_idx = (int32_t *)xcalloc(L->m, sizeof(int32_t));
  L->in_w = (double *)xcalloc(L->m, sizeof(double));
  int64_t *fill = (int64_t *)xcalloc(n, sizeof(int64_t));
  for (int i = 0; i < n; i++) fill[i] = L->in_ptr[i];
  for (int i = 0; i < n; i++)
    for (int64_t k = L->out_ptr[i]; k < L->out_ptr[i + 1]; k++) {
      int d = L->out_idx[k];
      L->in_idx[fill[d]] = i;
      L->in_w[fill[d]] = L->out_w[k];
      fill[d]++;
    }
  free(fill);
}

static void ensure_tab(orc_net *net, int deg) {
  int cap = 16;
  while (cap < 2 * deg + 4) cap *= 2;
  if (cap <= net->tab_cap) return;
  free(net->tab_key); free(net->tab_out); free(net->tab_in); free(net->tab_used); free(net->tab_order);
  net->tab_cap = cap;
  net->tab_key = (int32_t *)xcalloc(cap, sizeof(int32_t));
  for (int i = 0; i < cap; i++) net->tab_key[i] = -1;
  net->tab_out = (double *)xcalloc(cap, sizeof(double));
  net->tab_in = (double *)xcalloc(cap, sizeof(double));
  net->tab_used = (int32_t *)xcalloc(cap, sizeof(int32_t));
  net->tab_order = (int32_t *)xcalloc(cap, sizeof(int32_t));
}

/* Module::calibrate (Module.cpp:474-501) over modules 0..n_module-1. */
static void calibrate(orc_net *net) {
  double sel = 0.0, ssl = 0.0, se = 0.0;
  for (int i = 0; i < net->n_module; i++) {
    sel += plogp(net->m_exit[i]);
    ssl += plogp(net->m_stay[i]);
    se += net->m_exit[i];
  }
  net->sum_exit = se;
  net->codelength = plogp(se) - 2.0 * sel + ssl - net->all_log_all;
}

static void init_modules(orc_net *net);

int orc_initiate(orc_net *net) {
  level_t *L = &net->L;
  int n = net->n;
  L->dangling = (char *)xcalloc(n, 1);
  L->size = (double *)xcalloc(n, sizeof(double));
  L->dsize = (double *)xcalloc(n, sizeof(double));
  L->exitpr = (double *)xcalloc(n, sizeof(double));
  L->mod = (int32_t *)xcalloc(n, sizeof(int32_t));
  net->danglings = (int32_t *)xcalloc(n, sizeof(int32_t));
  net->n_dangling = 0;
  /* dangling detection + row normalisation (Module.cpp:220-249) */
  for (int i = 0; i < n; i++) {
    int64_t b = L->out_ptr[i], e = L->out_ptr[i + 1];
    if (b == e) { net->danglings[net->n_dangling++] = i; L->dangling[i] = 1; }
    else {
      double sum = 0.0;
      for (int64_t k = b; k < e; k++) sum += L->out_w[k];
      for (int64_t k = b; k < e; k++) L->out_w[k] /= sum;
    }
  }
  /* power iteration (Module.cpp:340-469) */
  double *tmp = (double *)xcalloc(n, sizeof(double));
  double *added = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; i++) tmp[i] = 1.0 / n;
  int iter = 0;
  double sqdiff = 1.0;
  do {
    double dsum = 0.0;
    for (int k = 0; k < net->n_dangling; k++) dsum += tmp[net->danglings[k]];
    for (int i = 0; i < n; i++) L->size[i] = (ALPHA + BETA * dsum) * L->tp[i];
    for (int i = 0; i < n; i++) added[i] = 0.0;
    for (int i = 0; i < n; i++)
      for (int64_t k = L->out_ptr[i]; k < L->out_ptr[i + 1]; k++)
        added[L->out_idx[k]] += BETA * L->out_w[k] * tmp[i];
    for (int i = 0; i < n; i++) L->size[i] += added[i];
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += L->size[i];
    sqdiff = 0.0;
    for (int i = 0; i < n; i++) {
      L->size[i] = L->size[i] / sum;
      sqdiff += fabs(L->size[i] - tmp[i]);
      tmp[i] = L->size[i];
    }
    iter++;
  } while (iter < 200 && (sqdiff > 1.0e-15 || iter < 50));
  free(tmp); free(added);
  net->iterations = iter;
  /* flows on arcs (Module.cpp:267-287) */
  for (int i = 0; i < n; i++) {
    if (!L->dangling[i]) {
      for (int64_t k = L->out_ptr[i]; k < L->out_ptr[i + 1]; k++) L->out_w[k] = L->size[i] * L->out_w[k];
    } else {
      L->dsize[i] = L->size[i];
    }
  }
  double all = 0.0;
  for (int i = 0; i < n; i++) all += plogp(L->size[i]);
  net->all_log_all = all;
  init_modules(net);
  return 0;
}

/* Singleton modules, Module::Module(int, Node*) (Module.cpp:67-99), then
 * calibrate, the in-links and the sweep state: the end of initiate. */
static void init_modules(orc_net *net) {
  level_t *L = &net->L;
  int n = net->n;
  /* singleton modules, Module::Module(int, Node*) (Module.cpp:67-99) */
  net->m_exit = (double *)xcalloc(n, sizeof(double));
  net->m_stay = (double *)xcalloc(n, sizeof(double));
  net->m_sum = (double *)xcalloc(n, sizeof(double));
  net->m_tp = (double *)xcalloc(n, sizeof(double));
  net->m_dang = (double *)xcalloc(n, sizeof(double));
  net->m_n = (int32_t *)xcalloc(n, sizeof(int32_t));
  net->mem_ptr = (int64_t *)xcalloc(n + 1, sizeof(int64_t));
  net->mem_idx = (int32_t *)xcalloc(n, sizeof(int32_t));
  net->mem_cnt = (int32_t *)xcalloc(n, sizeof(int32_t));
  net->empty = (int32_t *)xcalloc(n, sizeof(int32_t));
  int maxdeg = 0;
  for (int i = 0; i < n; i++) {
    double sumExit = 0.0;
    for (int64_t k = L->out_ptr[i]; k < L->out_ptr[i + 1]; k++) sumExit += L->out_w[k];
    double sp = L->size[i], ex;
    if (!L->dangling[i]) ex = ALPHA * (1.0 - L->tp[i]) * sp + BETA * sumExit;
    else { ex = (1.0 - L->tp[i]) * sp; net->m_dang[i] = L->size[i]; }
    L->exitpr[i] = ex;
    net->m_exit[i] = ex;
    net->m_sum[i] = sp;
    net->m_tp[i] = L->tp[i];
    net->m_stay[i] = ex + sp;
    net->m_n[i] = 1;
    net->mem_ptr[i + 1] = i + 1;
    net->mem_idx[i] = i;
    net->mem_cnt[i] = 1;
    L->mod[i] = i;
  }
  net->n_module = n;
  net->n_empty_mod = 0;
  net->n_empty_stack = 0;
  calibrate(net);
  build_in(L);
  for (int i = 0; i < n; i++) {
    int d = (int)(L->out_ptr[i + 1] - L->out_ptr[i] + L->in_ptr[i + 1] - L->in_ptr[i]);
    if (d > maxdeg) maxdeg = d;
  }
  ensure_tab(net, maxdeg + 1);
  net->is_active = (char *)xcalloc(n, 1);
  net->active = (int32_t *)xcalloc(n, sizeof(int32_t));
  mt_seed(&net->mt, 1u);
  orc_reset_active(net, 0);
}

/* generate_network_from_module (OmpRelaxmap.cpp:640-695): the network of
 * members[0..k) (in the order given: the reference lists a module's members
 * in ascending node id) with the parent's flows, teleport weights and
 * dangling flags and the parent's arc flows between members, in stored
 * order; then singleton modules over those arcs only and calibrate.  NULL
 * for an id out of range or repeated. */
orc_net *orc_subnetwork(const orc_net *parent, int k, const int32_t *members) {
  const level_t *P = &parent->L;
  if (k <= 0 || !P->size) return NULL;
  int32_t *newid = (int32_t *)xcalloc(parent->n, sizeof(int32_t));
  for (int i = 0; i < parent->n; i++) newid[i] = -1;
  for (int i = 0; i < k; i++) {
    if (members[i] < 0 || members[i] >= parent->n || newid[members[i]] >= 0) { free(newid); return NULL; }
    newid[members[i]] = i;
  }
  orc_net *net = net_alloc(k);
  level_t *L = &net->L;
  L->n = k;
  L->out_ptr = (int64_t *)xcalloc(k + 1, sizeof(int64_t));
  for (int i = 0; i < k; i++) {
    int a = members[i];
    for (int64_t e = P->out_ptr[a]; e < P->out_ptr[a + 1]; e++)
      if (newid[P->out_idx[e]] >= 0) L->out_ptr[i + 1]++;
  }
  for (int i = 0; i < k; i++) L->out_ptr[i + 1] += L->out_ptr[i];
  int64_t m = L->out_ptr[k];
  L->m = m;
  net->m = m;
  L->out_idx = (int32_t *)xcalloc(m > 0 ? m : 1, sizeof(int32_t));
  L->out_w = (double *)xcalloc(m > 0 ? m : 1, sizeof(double));
  net->raw_w = (double *)xcalloc(m > 0 ? m : 1, sizeof(double));
  int64_t o = 0;
  for (int i = 0; i < k; i++) {
    int a = members[i];
    for (int64_t e = P->out_ptr[a]; e < P->out_ptr[a + 1]; e++)
      if (newid[P->out_idx[e]] >= 0) {
        L->out_idx[o] = newid[P->out_idx[e]];
        L->out_w[o] = P->out_w[e];
        net->raw_w[o] = P->out_w[e];
        o++;
      }
  }
  L->tp = (double *)xcalloc(k, sizeof(double));
  L->size = (double *)xcalloc(k, sizeof(double));
  L->dsize = (double *)xcalloc(k, sizeof(double));
  L->exitpr = (double *)xcalloc(k, sizeof(double));
  L->dangling = (char *)xcalloc(k, 1);
  L->mod = (int32_t *)xcalloc(k, sizeof(int32_t));
  net->danglings = (int32_t *)xcalloc(k, sizeof(int32_t));
  double tot = 0.0;
  for (int i = 0; i < k; i++) {
    int a = members[i];
    free(net->names[i]);
    net->names[i] = parent->names && parent->names[a] ? strdup(parent->names[a]) : NULL;
    net->node_w[i] = parent->node_w[a];
    tot += parent->node_w[a];
    L->tp[i] = P->tp[a];
    L->size[i] = P->size[a];
    L->dsize[i] = P->dsize[a];
    L->dangling[i] = P->dangling[a];
    if (L->dangling[i]) net->danglings[net->n_dangling++] = i;
  }
  net->tot_node_w = tot;
  free(newid);
  double all = 0.0;
  for (int i = 0; i < k; i++) all += plogp(L->size[i]);
  net->all_log_all = all;
  init_modules(net);
  return net;
}

int orc_iterations(const orc_net *net) { return net->iterations; }
int orc_n_dangling(const orc_net *net) { return net->n_dangling; }
void orc_get_flows(const orc_net *net, double *p) { memcpy(p, net->L.size, sizeof(double) * net->n); }
void orc_get_node_exit(const orc_net *net, double *q) { memcpy(q, net->L.exitpr, sizeof(double) * net->n); }
void orc_get_teleport(const orc_net *net, double *tau) { memcpy(tau, net->L.tp, sizeof(double) * net->n); }
void orc_get_out_csr(const orc_net *net, int64_t *ptr, int32_t *idx, double *w) {
  memcpy(ptr, net->L.out_ptr, sizeof(int64_t) * (net->n + 1));
  memcpy(idx, net->L.out_idx, sizeof(int32_t) * net->m);
  memcpy(w, net->L.out_w, sizeof(double) * net->m);
}
void orc_get_raw_out_w(const orc_net *net, double *w) { memcpy(w, net->raw_w, sizeof(double) * net->m); }
double orc_codelength(const orc_net *net) { return net->codelength; }
double orc_sum_exit(const orc_net *net) { return net->sum_exit; }
double orc_all_log_all(const orc_net *net) { return net->all_log_all; }
int orc_n_module(const orc_net *net) { return net->n_module; }

/* ------------------------------------------------------------------ */
/* Members, partitions and codelength (Module.cpp:2881-3002)            */
/* ------------------------------------------------------------------ */
void orc_update_members(orc_net *net) {
  int n = net->n;
  for (int i = 0; i <= n; i++) net->mem_ptr[i] = 0;
  for (int i = 0; i < n; i++) net->mem_ptr[net->L.mod[i] + 1]++;
  for (int i = 0; i < n; i++) net->mem_ptr[i + 1] += net->mem_ptr[i];
  for (int i = 0; i < n; i++) net->mem_cnt[i] = (int32_t)(net->mem_ptr[i + 1] - net->mem_ptr[i]);
  int32_t *fill = (int32_t *)xcalloc(n, sizeof(int32_t));
  for (int i = 0; i < n; i++) {
    int m = net->L.mod[i];
    net->mem_idx[net->mem_ptr[m] + fill[m]++] = i;
  }
  free(fill);
}

int orc_set_partition(orc_net *net, const int32_t *part) {
  int n = net->n;
  for (int i = 0; i < n; i++) if (part[i] < 0 || part[i] >= n) return -1;
  for (int m = 0; m < n; m++) {
    net->m_sum[m] = net->m_tp[m] = net->m_dang[m] = net->m_exit[m] = net->m_stay[m] = 0.0;
    net->m_n[m] = 0;
  }
  int nmod = 0;
  for (int i = 0; i < n; i++) {
    int m = part[i];
    net->L.mod[i] = m;
    if (net->m_n[m] == 0) nmod++;
    net->m_n[m]++;
    net->m_sum[m] += net->L.size[i];
    net->m_tp[m] += net->L.tp[i];
    if (net->L.dangling[i]) net->m_dang[m] += net->L.size[i];
  }
  net->n_empty_stack = 0;
  for (int m = 0; m < n; m++) if (net->m_n[m] == 0) net->empty[net->n_empty_stack++] = m;
  net->n_module = nmod;
  net->n_empty_mod = n - nmod;
  orc_update_members(net);
  /* Network::updateCodeLength(numTh, false) (Module.cpp:2918-2962) */
  double se = 0.0, el = 0.0, sl = 0.0;
  for (int m = 0; m < n; m++) {
    if (net->m_n[m] > 0) {
      double ex = ALPHA * (1.0 - net->m_tp[m]) * net->m_sum[m];
      double out = 0.0;
      for (int64_t j = net->mem_ptr[m]; j < net->mem_ptr[m + 1]; j++) {
        int a = net->mem_idx[j];
        for (int64_t k = net->L.out_ptr[a]; k < net->L.out_ptr[a + 1]; k++)
          if (net->L.mod[net->L.out_idx[k]] != m) out += net->L.out_w[k];
      }
      ex += BETA * (out + (1.0 - net->m_tp[m]) * net->m_dang[m]);
      net->m_exit[m] = ex;
      net->m_stay[m] = ex + net->m_sum[m];
      se += ex;
      el += plogp(ex);
      sl += plogp(net->m_stay[m]);
    }
  }
  net->sum_exit = se;
  net->codelength = plogp(se) - 2.0 * el + sl - net->all_log_all;
  net->has_super = 0;
  return 0;
}

void orc_get_partition(const orc_net *net, int32_t *part) { memcpy(part, net->L.mod, sizeof(int32_t) * net->n); }

void orc_get_module_stats(const orc_net *net, double *exit, double *sum_pr, double *sum_tp,
                          double *sum_dangling, int32_t *n_members) {
  int n = net->n;
  if (exit) memcpy(exit, net->m_exit, sizeof(double) * n);
  if (sum_pr) memcpy(sum_pr, net->m_sum, sizeof(double) * n);
  if (sum_tp) memcpy(sum_tp, net->m_tp, sizeof(double) * n);
  if (sum_dangling) memcpy(sum_dangling, net->m_dang, sizeof(double) * n);
  if (n_members) memcpy(n_members, net->m_n, sizeof(int32_t) * n);
}

double orc_calculate_codelength(const orc_net *net) {
  int n = net->n;
  double se = 0.0, el = 0.0, sl = 0.0;
  for (int m = 0; m < n; m++) {
    if (net->m_n[m] > 0) {
      double ex = ALPHA * (1.0 - net->m_tp[m]) * net->m_sum[m];
      double out = 0.0;
      for (int64_t j = net->mem_ptr[m]; j < net->mem_ptr[m + 1]; j++) {
        int a = net->mem_idx[j];
        for (int64_t k = net->L.out_ptr[a]; k < net->L.out_ptr[a + 1]; k++)
          if (net->L.mod[net->L.out_idx[k]] != m) out += net->L.out_w[k];
      }
      ex += BETA * (out + (1.0 - net->m_tp[m]) * net->m_dang[m]);
      se += ex;
      el += plogp(ex);
      sl += plogp(ex + net->m_sum[m]);
    }
  }
  return plogp(se) - 2.0 * el + sl - net->all_log_all;
}

/* ------------------------------------------------------------------ */
/* Contraction (Module.cpp:3008-3077, Node.cpp:76-106)                  */
/* ------------------------------------------------------------------ */
typedef struct { int32_t key; double w; } kw_t;
static int kw_cmp(const void *x, const void *y) {
  int32_t a = ((const kw_t *)x)->key, b = ((const kw_t *)y)->key;
  return (a > b) - (a < b);
}

int orc_contract(orc_net *net) {
  int n = net->n;
  level_t *L = &net->L, *S = &net->S;
  level_free(S);
  int32_t *mod2sp = (int32_t *)xcalloc(n, sizeof(int32_t));
  int ns = 0;
  for (int m = 0; m < n; m++) if (net->m_n[m] > 0) mod2sp[m] = ns++;
  S->n = ns;
  S->size = (double *)xcalloc(ns, sizeof(double));
  S->exitpr = (double *)xcalloc(ns, sizeof(double));
  S->tp = (double *)xcalloc(ns, sizeof(double));
  S->dsize = (double *)xcalloc(ns, sizeof(double));
  S->dangling = (char *)xcalloc(ns, 1);
  S->mod = (int32_t *)xcalloc(ns, sizeof(int32_t));
  S->mem_ptr = (int64_t *)xcalloc(ns + 1, sizeof(int64_t));
  S->mem_idx = (int32_t *)xcalloc(n, sizeof(int32_t));
  int s = 0;
  for (int m = 0; m < n; m++) {
    if (net->m_n[m] > 0) {
      S->size[s] = net->m_sum[m];
      S->exitpr[s] = net->m_exit[m];
      S->tp[s] = net->m_tp[m];
      S->dsize[s] = net->m_dang[m];
      S->mod[s] = m;
      int64_t cnt = net->mem_ptr[m + 1] - net->mem_ptr[m];
      S->mem_ptr[s + 1] = S->mem_ptr[s] + cnt;
      memcpy(S->mem_idx + S->mem_ptr[s], net->mem_idx + net->mem_ptr[m], sizeof(int32_t) * cnt);
      s++;
    }
  }
  /* inter-module arcs, summed per (super, super) in member / link order;
   * std::map iteration gives ascending destination order */
  S->out_ptr = (int64_t *)xcalloc(ns + 1, sizeof(int64_t));
  int64_t cap = 1024, total = 0;
  int32_t *oi = (int32_t *)xcalloc(cap, sizeof(int32_t));
  double *ow = (double *)xcalloc(cap, sizeof(double));
  int32_t *slot = (int32_t *)xcalloc(ns, sizeof(int32_t));
  for (int i = 0; i < ns; i++) slot[i] = -1;
  kw_t *buf = NULL;
  int64_t bufcap = 0;
  for (s = 0; s < ns; s++) {
    int64_t nb = 0;
    for (int64_t j = S->mem_ptr[s]; j < S->mem_ptr[s + 1]; j++) {
      int a = S->mem_idx[j];
      for (int64_t k = L->out_ptr[a]; k < L->out_ptr[a + 1]; k++) {
        int t = mod2sp[L->mod[L->out_idx[k]]];
        if (t == s) continue;
        if (slot[t] < 0) {
          if (nb == bufcap) { bufcap = bufcap ? 2 * bufcap : 64; buf = (kw_t *)realloc(buf, sizeof(kw_t) * bufcap); }
          slot[t] = (int32_t)nb;
          buf[nb].key = t;
          buf[nb].w = L->out_w[k];
          nb++;
        } else {
          buf[slot[t]].w += L->out_w[k];
        }
      }
    }
    for (int64_t j = 0; j < nb; j++) slot[buf[j].key] = -1;
    qsort(buf, (size_t)nb, sizeof(kw_t), kw_cmp);
    while (total + nb > cap) {
      cap *= 2;
      oi = (int32_t *)realloc(oi, sizeof(int32_t) * cap);
      ow = (double *)realloc(ow, sizeof(double) * cap);
    }
    for (int64_t j = 0; j < nb; j++) { oi[total + j] = buf[j].key; ow[total + j] = buf[j].w; }
    total += nb;
    S->out_ptr[s + 1] = total;
  }
  free(buf); free(slot); free(mod2sp);
  S->out_idx = oi;
  S->out_w = ow;
  S->m = total;
  build_in(S);
  net->has_super = 1;
  int maxdeg = 0;
  for (int i = 0; i < ns; i++) {
    int d = (int)(S->out_ptr[i + 1] - S->out_ptr[i] + S->in_ptr[i + 1] - S->in_ptr[i]);
    if (d > maxdeg) maxdeg = d;
  }
  ensure_tab(net, maxdeg + 1);
  return ns;
}

int orc_n_super(const orc_net *net) { return net->has_super ? net->S.n : 0; }
int64_t orc_super_m(const orc_net *net) { return net->has_super ? net->S.m : 0; }
void orc_get_super(const orc_net *net, double *size, double *exit, double *tp, double *dangling,
                   int32_t *mod, int32_t *n_members) {
  const level_t *S = &net->S;
  for (int i = 0; i < S->n; i++) {
    if (size) size[i] = S->size[i];
    if (exit) exit[i] = S->exitpr[i];
    if (tp) tp[i] = S->tp[i];
    if (dangling) dangling[i] = S->dsize[i];
    if (mod) mod[i] = S->mod[i];
    if (n_members) n_members[i] = (int32_t)(S->mem_ptr[i + 1] - S->mem_ptr[i]);
  }
}
void orc_get_super_out(const orc_net *net, int64_t *ptr, int32_t *idx, double *w) {
  const level_t *S = &net->S;
  memcpy(ptr, S->out_ptr, sizeof(int64_t) * (S->n + 1));
  memcpy(idx, S->out_idx, sizeof(int32_t) * S->m);
  memcpy(w, S->out_w, sizeof(double) * S->m);
}
void orc_get_super_in(const orc_net *net, int64_t *ptr, int32_t *idx, double *w) {
  const level_t *S = &net->S;
  memcpy(ptr, S->in_ptr, sizeof(int64_t) * (S->n + 1));
  memcpy(idx, S->in_idx, sizeof(int32_t) * S->m);
  memcpy(w, S->in_w, sizeof(double) * S->m);
}

/* ------------------------------------------------------------------ */
/* Move sweeps (Module.cpp:511-2869)                                    */
/* ------------------------------------------------------------------ */
void orc_seed(orc_net *net, uint32_t seed) { mt_seed(&net->mt, seed); }
uint32_t orc_rand_int(orc_net *net, uint32_t n) { return mt_rand_n(&net->mt, n); }

void orc_reset_active(orc_net *net, int sp) {
  int n = sp ? net->S.n : net->n;
  for (int i = 0; i < n; i++) { net->active[i] = i; net->is_active[i] = 0; }
  net->n_active = n;
}
int orc_n_active(const orc_net *net) { return net->n_active; }

/* neighbour-module table: insertion keeps per-key link order */
static int tab_slot(orc_net *net, int key, int *nused) {
  uint32_t mask = (uint32_t)net->tab_cap - 1u;
  uint32_t h = ((uint32_t)key * 2654435761u) & mask;
  while (net->tab_key[h] != -1 && net->tab_key[h] != key) h = (h + 1) & mask;
  if (net->tab_key[h] == -1) {
    net->tab_key[h] = key;
    net->tab_out[h] = 0.0;
    net->tab_in[h] = 0.0;
    net->tab_used[(*nused)++] = (int32_t)h;
  }
  return (int)h;
}

static int32_t *g_sort_keys;
static int ord_cmp(const void *x, const void *y) {
  int32_t a = g_sort_keys[*(const int32_t *)x], b = g_sort_keys[*(const int32_t *)y];
  return (a > b) - (a < b);
}

typedef struct { double diff; int mod; double sum1, sum2, exit1, exit2, newsum; } move_t;

/* Evaluates unit u of level V (Module.cpp:531-662 and its twins).  On
 * return the table holds the candidate keys; *best has diff < 0 if any
 * candidate improves.  variant >= 2 uses the parallel sentinel for the
 * empty module (Module.cpp:986) whose statistics the reference reads past
 * the end of its module array; this restatement reads an empty module
 * (all zero) there. */
static void evaluate(orc_net *net, level_t *V, int sp, int variant, int u, move_t *best,
                     double *out_old, double *in_old) {
  int nused = 0;
  int oldMod = V->mod[u];
  const double beta = BETA, alpha = ALPHA;
  for (int64_t k = V->out_ptr[u]; k < V->out_ptr[u + 1]; k++) {
    int h = tab_slot(net, V->mod[V->out_idx[k]], &nused);
    net->tab_out[h] += beta * V->out_w[k];
  }
  for (int64_t k = V->in_ptr[u]; k < V->in_ptr[u + 1]; k++) {
    int h = tab_slot(net, V->mod[V->in_idx[k]], &nused);
    net->tab_in[h] += beta * V->in_w[k];
  }
  double ndSize = V->size[u], ndTP = V->tp[u], ndD = V->dsize[u], ndExit = V->exitpr[u];
  double oldExit1 = net->m_exit[oldMod], oldSum1 = net->m_sum[oldMod];
  double oldD1 = net->m_dang[oldMod], oldTP = net->m_tp[oldMod];
  double addOut = (alpha * ndSize + beta * ndD) * (oldTP - ndTP);
  double addIn = (alpha * (oldSum1 - ndSize) + beta * (oldD1 - ndD)) * ndTP;
  /* keys in ascending order */
  for (int j = 0; j < nused; j++) net->tab_order[j] = net->tab_used[j];
  g_sort_keys = net->tab_key;
  qsort(net->tab_order, (size_t)nused, sizeof(int32_t), ord_cmp);
  int has_old = 0;
  for (int j = 0; j < nused; j++) {
    int h = net->tab_order[j], m = net->tab_key[h];
    if (m == oldMod) { net->tab_out[h] += addOut; net->tab_in[h] += addIn; has_old = 1; *out_old = net->tab_out[h]; *in_old = net->tab_in[h]; }
    else {
      net->tab_out[h] += (alpha * ndSize + beta * ndD) * net->m_tp[m];
      net->tab_in[h] += (alpha * net->m_sum[m] + beta * net->m_dang[m]) * ndTP;
    }
  }
  if (!has_old) { *out_old = addOut; *in_old = addIn; }
  /* empty-module option (Module.cpp:606-611, 1073-1077, 1791-1796, 2263-2267) */
  int offer;
  if (!sp) offer = net->mem_cnt[oldMod] > 1;
  else if (variant < 2) offer = (double)net->mem_cnt[oldMod] > ndSize;
  else offer = net->m_sum[oldMod] > ndSize;
  offer = offer && net->n_empty_stack > 0;
  int empty_key = -1;
  if (offer) {
    empty_key = variant < 2 ? net->empty[net->n_empty_stack - 1] : net->n + 1;
    int before = nused;
    int h = tab_slot(net, empty_key, &nused);
    net->tab_out[h] = 0.0;
    net->tab_in[h] = 0.0;
    if (nused > before) {
      net->tab_order[nused - 1] = h;
      qsort(net->tab_order, (size_t)nused, sizeof(int32_t), ord_cmp);
    }
  }
  double newExit1 = oldExit1 - ndExit + *out_old + *in_old;
  best->diff = 0.0;
  best->mod = -1;
  double S = net->sum_exit;
  for (int j = 0; j < nused; j++) {
    int h = net->tab_order[j], m = net->tab_key[h];
    if (m == oldMod) continue;
    if (net->eval_units && m > oldMod && m < net->n && net->eval_units[oldMod] == 1 && net->eval_units[m] == 1)
      continue;   /* the batch rule of orc_evaluate_all (no singleton joins a larger singleton) */
    double outNew = net->tab_out[h], inNew = net->tab_in[h];
    int real = m < net->n;
    double oldExit2 = real ? net->m_exit[m] : 0.0, oldSum2 = real ? net->m_sum[m] : 0.0;
    move_t c;
    c.mod = m;
    c.sum1 = oldSum1 - ndSize;
    c.sum2 = oldSum2 + ndSize;
    c.exit1 = newExit1;
    c.exit2 = oldExit2 + ndExit - outNew - inNew;
    c.newsum = S + newExit1 + c.exit2 - oldExit1 - oldExit2;
    double dAll = plogp(c.newsum) - plogp(S);
    double dExit = plogp(c.exit1) + plogp(c.exit2) - plogp(oldExit1) - plogp(oldExit2);
    double dStay = plogp(c.exit1 + c.sum1) + plogp(c.exit2 + c.sum2) - plogp(oldExit1 + oldSum1) - plogp(oldExit2 + oldSum2);
    c.diff = dAll - 2.0 * dExit + dStay;
    if (net->cand_mod && net->cand_n < net->cand_cap) {
      net->cand_mod[net->cand_n] = m;
      net->cand_diff[net->cand_n] = c.diff;
      net->cand_n++;
    }
    if (c.diff < best->diff) *best = c;
  }
  for (int j = 0; j < nused; j++) net->tab_key[net->tab_used[j]] = -1;
}

/* Every unit of the level evaluated against the same module statistics, as
 * prioritize_parallelMove's decision loop does for each node before its
 * commit (Module.cpp:1355-1488, and the SP twin Module.cpp:2154-2320):
 * candidates in ascending module id with the empty-module sentinel n + 1
 * last (Module.cpp:1348, 1435-1439), strict < from a best of 0.0
 * (Module.cpp:1451, 1484).  best_mod[u] = -1 when no candidate has dL < 0.
 * no_singleton_swaps applies a rule that is not the reference's: a unit
 * alone in its module skips a candidate module of larger id that also holds
 * one unit (the GPU's batch rule, which stands in for the commit-time check
 * of Module.cpp:1508-1515 that rejects the second move of a swap). */
void orc_evaluate_all(orc_net *net, int sp, int no_singleton_swaps, int32_t *best_mod, double *best_diff) {
  level_t *V = sp ? &net->S : &net->L;
  int32_t *units = NULL;
  if (no_singleton_swaps) {
    units = (int32_t *)xcalloc((size_t)net->n + 2, sizeof(int32_t));
    for (int u = 0; u < V->n; u++) units[V->mod[u]]++;
  }
  net->eval_units = units;
  for (int u = 0; u < V->n; u++) {
    move_t best;
    double oo, io;
    evaluate(net, V, sp, 3, u, &best, &oo, &io);
    best_mod[u] = best.diff < 0.0 ? best.mod : -1;
    best_diff[u] = best.diff < 0.0 ? best.diff : 0.0;
  }
  net->eval_units = NULL;
  free(units);
}

/* The candidates of unit u with their dL, in the order evaluated (ascending
 * module id, the empty-module sentinel n + 1 last), the unit's own module
 * left out.  Returns the number written (at most cap). */
int orc_candidates(orc_net *net, int sp, int u, int cap, int32_t *mods, double *diffs) {
  level_t *V = sp ? &net->S : &net->L;
  if (u < 0 || u >= V->n) return -1;
  net->cand_mod = mods;
  net->cand_diff = diffs;
  net->cand_n = 0;
  net->cand_cap = cap;
  move_t best;
  double oo, io;
  evaluate(net, V, sp, 3, u, &best, &oo, &io);
  int k = net->cand_n;
  net->cand_mod = NULL;
  net->cand_diff = NULL;
  net->cand_n = net->cand_cap = 0;
  return k;
}

/* Applies a move of unit u to newMod with precomputed statistics
 * (Module.cpp:669-706 and its twins). */
static void apply_move(orc_net *net, level_t *V, int sp, int u, int newMod, const move_t *r, int parallel) {
  int oldMod = V->mod[u];
  double ndTP = V->tp[u], ndSize = V->size[u], ndD = V->dsize[u];
  int cnt = 1;
  V->mod[u] = newMod;
  if (sp) {
    cnt = (int)(V->mem_ptr[u + 1] - V->mem_ptr[u]);
    for (int64_t j = V->mem_ptr[u]; j < V->mem_ptr[u + 1]; j++) net->L.mod[V->mem_idx[j]] = newMod;
  }
  net->m_n[newMod] += cnt;
  net->m_exit[newMod] = r->exit2;
  net->m_sum[newMod] = r->sum2;
  net->m_stay[newMod] = r->exit2 + r->sum2;
  net->m_tp[newMod] += ndTP;
  if (sp ? ndD > 0.0 : V->dangling[u] != 0) {
    double d = (sp || parallel) ? ndD : ndSize;
    net->m_dang[newMod] += d;
    net->m_dang[oldMod] -= d;
  }
  net->m_n[oldMod] -= cnt;
  net->m_exit[oldMod] = r->exit1;
  net->m_sum[oldMod] = r->sum1;
  net->m_stay[oldMod] = r->exit1 + r->sum1;
  net->m_tp[oldMod] -= ndTP;
  if (net->m_n[oldMod] == 0) {
    net->n_empty_mod++;
    net->n_module--;
    net->empty[net->n_empty_stack++] = oldMod;
  }
  net->sum_exit = r->newsum;
  net->codelength += r->diff;
}

/* The commit of the parallel variants (Module.cpp:1133-1295): exact
 * re-evaluation against current module statistics, then an unconditional
 * move.  Returns 1 if the move was made. */
static int parallel_commit(orc_net *net, level_t *V, int sp, int u, int newMod) {
  int oldMod = V->mod[u];
  const int emptyTarget = net->n + 1;
  int isEmpty = 0;
  const double alpha = ALPHA, beta = BETA;
  if (net->n_empty_mod > 0 && newMod == emptyTarget && net->m_n[oldMod] > 1) {
    newMod = net->empty[net->n_empty_stack - 1];
    isEmpty = 1;
  } else if (newMod == emptyTarget) return 0;
  else if (net->m_n[newMod] == 0) return 0;
  double outOld = 0.0, outNew = 0.0, inOld = 0.0, inNew = 0.0;
  for (int64_t k = V->out_ptr[u]; k < V->out_ptr[u + 1]; k++) {
    int t = V->mod[V->out_idx[k]];
    if (t == oldMod) outOld += beta * V->out_w[k];
    else if (!isEmpty && t == newMod) outNew += beta * V->out_w[k];
  }
  for (int64_t k = V->in_ptr[u]; k < V->in_ptr[u + 1]; k++) {
    int t = V->mod[V->in_idx[k]];
    if (t == oldMod) inOld += beta * V->in_w[k];
    else if (!isEmpty && t == newMod) inNew += beta * V->in_w[k];
  }
  double ndSize = V->size[u], ndTP = V->tp[u], ndD = V->dsize[u], ndExit = V->exitpr[u];
  double oldExit1 = net->m_exit[oldMod], oldSum1 = net->m_sum[oldMod];
  double oldD1 = net->m_dang[oldMod], oldTP = net->m_tp[oldMod];
  outOld += (alpha * ndSize + beta * ndD) * (oldTP - ndTP);
  inOld += (alpha * (oldSum1 - ndSize) + beta * (oldD1 - ndD)) * ndTP;
  outNew += (alpha * ndSize + beta * ndD) * net->m_tp[newMod];
  inNew += (alpha * net->m_sum[newMod] + beta * net->m_dang[newMod]) * ndTP;
  if (isEmpty) { outNew = 0.0; inNew = 0.0; }
  move_t r;
  r.exit1 = oldExit1 - ndExit + outOld + inOld;
  double oldExit2 = net->m_exit[newMod], oldSum2 = net->m_sum[newMod];
  r.mod = newMod;
  r.sum1 = oldSum1 - ndSize;
  r.sum2 = oldSum2 + ndSize;
  r.exit2 = oldExit2 + ndExit - outNew - inNew;
  double S = net->sum_exit;
  r.newsum = S + r.exit1 + r.exit2 - oldExit1 - oldExit2;
  double dAll = plogp(r.newsum) - plogp(S);
  double dExit = plogp(r.exit1) + plogp(r.exit2) - plogp(oldExit1) - plogp(oldExit2);
  double dStay = plogp(r.exit1 + r.sum1) + plogp(r.exit2 + r.sum2) - plogp(oldExit1 + oldSum1) - plogp(oldExit2 + oldSum2);
  r.diff = dAll - 2.0 * dExit + dStay;
  if (isEmpty) {
    net->n_empty_stack--;
    net->n_empty_mod--;
    net->n_module++;
  }
  apply_move(net, V, sp, u, newMod, &r, 1);
  return 1;
}

int orc_sweep(orc_net *net, int variant, int sp, double vthresh) {
  level_t *V = sp ? &net->S : &net->L;
  int prior = variant == 1 || variant == 3;
  int parallel = variant >= 2;
  int n = prior ? net->n_active : V->n;
  int32_t *order = (int32_t *)xcalloc(n, sizeof(int32_t));
  for (int i = 0; i < n; i++) order[i] = prior ? net->active[i] : i;
  for (int i = 0; i < n; i++) {
    uint32_t t = mt_rand_n(&net->mt, (uint32_t)(n - 1));
    int32_t x = order[i]; order[i] = order[t]; order[t] = x;
  }
  double thr = prior ? vthresh : 0.0;
  int moved = 0;
  for (int i = 0; i < n; i++) {
    int u = order[i];
    move_t best;
    double oo, io;
    evaluate(net, V, sp, variant, u, &best, &oo, &io);
    if (!(best.diff < thr)) continue;
    if (!parallel) {
      int newMod = best.mod;
      if (net->m_n[newMod] == 0) {
        newMod = net->empty[--net->n_empty_stack];
        net->n_empty_mod--;
        net->n_module++;
      }
      int cnt = sp ? (int)(V->mem_ptr[u + 1] - V->mem_ptr[u]) : 1;
      apply_move(net, V, sp, u, newMod, &best, 0);
      moved += cnt;
    } else {
      int cnt = sp ? (int)(V->mem_ptr[u + 1] - V->mem_ptr[u]) : 1;
      if (parallel_commit(net, V, sp, u, best.mod)) moved += cnt;
    }
    if (prior) {
      for (int64_t k = V->out_ptr[u]; k < V->out_ptr[u + 1]; k++) net->is_active[V->out_idx[k]] = 1;
      for (int64_t k = V->in_ptr[u]; k < V->in_ptr[u + 1]; k++) net->is_active[V->in_idx[k]] = 1;
    }
  }
  free(order);
  if (prior) {
    int k = 0;
    for (int i = 0; i < V->n; i++)
      if (net->is_active[i]) { net->active[k++] = i; net->is_active[i] = 0; }
    net->n_active = k;
  }
  return moved;
}

/* ------------------------------------------------------------------ */
/* Driver (OmpRelaxmap.cpp:181-219, 263-429) at numTh == 1              */
/* ------------------------------------------------------------------ */
static void trace_push(orc_net *net, int moved) {
  if (net->n_trace == net->trace_cap) {
    net->trace_cap = net->trace_cap ? 2 * net->trace_cap : 64;
    net->trace = (int32_t *)realloc(net->trace, sizeof(int32_t) * net->trace_cap);
  }
  net->trace[net->n_trace++] = moved;
}

int orc_trace_len(const orc_net *net) { return net->n_trace; }
void orc_get_trace(const orc_net *net, int32_t *moved) { memcpy(moved, net->trace, sizeof(int32_t) * net->n_trace); }
int orc_n_empty(const orc_net *net) { return net->n_empty_mod; }

static void greedy_partition(orc_net *net, double threshold, double vthresh, int max_iter,
                             int prior, int fine) {
  const double ln2 = log(2.0);
  int iter = 0, stop = 0;
  orc_reset_active(net, !fine);
  while (!stop && iter < max_iter) {
    double old = net->codelength;
    trace_push(net, orc_sweep(net, prior ? 1 : 0, !fine, vthresh));
    iter++;
    if (old - net->codelength >= 0 && (old - net->codelength) / ln2 < threshold) stop = 1;
  }
  orc_update_members(net);
  double old;
  do {
    old = net->codelength;
    stop = 0;
    orc_contract(net);
    orc_reset_active(net, 1);
    int sp_iter = 0;
    while (!stop && sp_iter < max_iter) {
      double inner = net->codelength;
      trace_push(net, orc_sweep(net, prior ? 1 : 0, 1, vthresh));
      sp_iter++;
      if (inner - net->codelength >= 0.0 && (inner - net->codelength) / ln2 < threshold) stop = 1;
    }
    orc_update_members(net);
  } while ((old - net->codelength) / ln2 > threshold);
}

int orc_run(orc_net *net, double threshold, double vthresh_neg, int max_iter, int prior,
            int *coarse_next) {
  const double ln2 = log(2.0);
  int fine = 1, steps = 0;
  *coarse_next = 0;
  double old = net->codelength;
  greedy_partition(net, threshold, vthresh_neg, max_iter, prior, fine);
  steps = 1;
  int next = !((old - net->codelength) / ln2 < threshold);
  while (next) {
    old = net->codelength;
    greedy_partition(net, threshold, vthresh_neg, max_iter, prior, fine);
    steps++;
    if ((old - net->codelength) / ln2 < threshold) next = 0;
    fine = !fine;
    if (next && !fine) { *coarse_next = 1; break; }
  }
  return steps;
}
