// Seeded synthetic networks for the benchmark and the tests (include/relaxmap_synth.h).
//
// Every node (or, for R-MAT, every sampled arc) owns a counter-based random
// stream (splitmix64 of (seed, stream id)), so the output does not depend on
// the number of OpenMP threads.  Node ids of the community models are a
// pseudo-random permutation of community-sorted positions (a 4-round Feistel
// bijection with cycle walking), so communities are not contiguous id ranges.
#include "../../include/relaxmap_synth.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace {
// An allocator whose value construction leaves the memory uninitialised, so
// that resizing the multi-GB arc arrays does not zero them on one thread (the
// parallel loops that fill them touch their pages first).
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U> &) {}
  template <typename U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <typename T>
using big_vector = std::vector<T, NoInit<T>>;
}  // namespace

struct rmx_syn {
  int32_t n = 0;
  std::vector<int64_t> ptr;
  big_vector<int32_t> idx;
  big_vector<double> w;
  std::vector<int32_t> com;
};

namespace {

inline uint64_t smix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t stream, uint64_t salt) : s(smix(seed ^ smix(stream * 0xd1342543de82ef95ull + salt))) {}
  uint64_t next() { return smix(s += 0x9e3779b97f4a7c15ull); }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
  double uniform_pos() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }   // (0,1]
  uint64_t below(uint64_t n) { return n ? (uint64_t)(((__uint128_t)next() * n) >> 64) : 0; }
  double normal() {   // Box-Muller
    double u1 = uniform_pos(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int poisson(double lam) {   // Knuth for small means, normal approximation above 60
    if (lam > 60.0) {
      double v = lam + std::sqrt(lam) * normal();
      return v < 0 ? 0 : (int)std::lround(v);
    }
    double L = std::exp(-lam), p = 1.0;
    int k = 0;
    do {
      ++k;
      p *= uniform();
    } while (p > L);
    return k - 1;
  }
};

// continuous power law p(x) ~ x^-g on [a, b]
double power_law(Rng &r, double g, double a, double b) {
  double u = r.uniform();
  if (std::fabs(g - 1.0) < 1e-12) return a * std::pow(b / a, u);
  double e = 1.0 - g;
  double lo = std::pow(a, e), hi = std::pow(b, e);
  return std::pow(lo + u * (hi - lo), 1.0 / e);
}

// bijection on [0, n): Feistel network on the next even power of two, cycle walking
struct Perm {
  uint64_t n, half_bits, mask;
  uint64_t keys[4];
  Perm(uint64_t n_, uint64_t seed) : n(n_) {
    int bits = 2;
    while ((1ull << bits) < n) bits += 2;
    half_bits = bits / 2;
    mask = (1ull << half_bits) - 1;
    for (int i = 0; i < 4; ++i) keys[i] = smix(seed * 4 + i + 0x1234);
  }
  uint64_t round_f(uint64_t x, int i) const { return smix(x ^ keys[i]) & mask; }
  uint64_t enc1(uint64_t x) const {
    uint64_t l = x >> half_bits, r = x & mask;
    for (int i = 0; i < 4; ++i) {
      uint64_t t = r;
      r = l ^ round_f(r, i);
      l = t;
    }
    return (l << half_bits) | r;
  }
  uint64_t dec1(uint64_t x) const {
    uint64_t l = x >> half_bits, r = x & mask;
    for (int i = 3; i >= 0; --i) {
      uint64_t t = l;
      l = r ^ round_f(l, i);
      r = t;
    }
    return (l << half_bits) | r;
  }
  uint64_t fwd(uint64_t x) const {
    do { x = enc1(x); } while (x >= n);
    return x;
  }
  uint64_t inv(uint64_t x) const {
    do { x = dec1(x); } while (x >= n);
    return x;
  }
};

// community sizes summing to n (the last one absorbed if below the minimum)
std::vector<int64_t> community_bounds(int32_t n, double g, int32_t cmin, int32_t cmax, uint64_t seed) {
  std::vector<int64_t> b{0};
  Rng r(seed, 0xC0FFEE, 7);
  int64_t tot = 0;
  while (tot < n) {
    int64_t s = (int64_t)std::floor(power_law(r, g, cmin, cmax + 1.0));
    s = std::max<int64_t>(cmin, std::min<int64_t>(cmax, s));
    if (tot + s > n) s = n - tot;
    if (s < cmin && b.size() > 1) {
      b.back() = n;
      tot = n;
      break;
    }
    tot += s;
    b.push_back(tot);
  }
  return b;
}

// per-node arc lists -> CSR (sorted, self-loops dropped, duplicates summed in
// arc order)
struct Builder {
  int32_t n;
  std::vector<int64_t> off;   // raw offsets (before dedup)
  big_vector<int32_t> tgt;
  big_vector<double> wt;
  explicit Builder(int32_t n_) : n(n_), off((size_t)n_ + 1, 0) {}
  void alloc() {
    for (int32_t i = 0; i < n; ++i) off[i + 1] += off[i];
    tgt.resize(off[n]);
    wt.resize(off[n]);
  }
  void finish(rmx_syn &g) {
    std::vector<int64_t> deg((size_t)n + 1, 0);
#pragma omp parallel
    {
      std::vector<std::pair<int32_t, int64_t>> tmp;   // per thread, reused across nodes
      std::vector<double> ws;
#pragma omp for schedule(dynamic, 1024)
      for (int32_t u = 0; u < n; ++u) {
        int64_t b = off[u], e = off[u + 1];
        // sort by target, stable in arc order so duplicate weights sum in order
        tmp.clear();
        for (int64_t k = b; k < e; ++k)
          if (tgt[k] != u) tmp.push_back({tgt[k], k});
        std::sort(tmp.begin(), tmp.end());
        ws.clear();
        int64_t o = b;
        for (size_t i = 0; i < tmp.size(); ++i) {
          double wv = wt[tmp[i].second];
          if (i > 0 && tmp[i].first == tmp[i - 1].first) {
            ws.back() += wv;
          } else {
            tgt[o] = tmp[i].first;
            ws.push_back(wv);
            ++o;
          }
        }
        for (size_t i = 0; i < ws.size(); ++i) wt[b + i] = ws[i];
        deg[u + 1] = o - b;
      }
    }
    g.n = n;
    g.ptr.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + deg[i + 1];
    g.idx.resize(g.ptr[n]);
    g.w.resize(g.ptr[n]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t u = 0; u < n; ++u) {
      int64_t d = g.ptr[u + 1] - g.ptr[u];
      std::memcpy(&g.idx[g.ptr[u]], &tgt[off[u]], sizeof(int32_t) * d);
      std::memcpy(&g.w[g.ptr[u]], &wt[off[u]], sizeof(double) * d);
    }
    big_vector<int32_t>().swap(tgt);
    big_vector<double>().swap(wt);
  }
};

int community_model(int32_t n, double deg_exp, int32_t deg_min, int32_t deg_max, double mean_deg, double com_exp,
                    int32_t com_min, int32_t com_max, double mu, double dangling, uint64_t seed, bool chung_lu,
                    rmx_syn **out) {
  if (n < 2 || deg_min < 1 || deg_max < deg_min || com_min < 1 || com_max < com_min || mu < 0 || mu > 1)
    return -22;
  rmx_syn *g = new rmx_syn();
  const std::vector<int64_t> cb = community_bounds(n, com_exp, com_min, com_max, seed);
  const int64_t ncom = (int64_t)cb.size() - 1;
  std::vector<int32_t> com_of_pos(n);
  for (int64_t c = 0; c < ncom; ++c)
    for (int64_t p = cb[c]; p < cb[c + 1]; ++p) com_of_pos[p] = (int32_t)c;
  const Perm perm((uint64_t)n, seed ^ 0xABCDEFull);
  // out-degree targets (raw power law samples) and in-weights by position
  std::vector<double> raw(n), win;
  if (chung_lu) win.resize(n);
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    Rng r(seed, (uint64_t)p, 1);
    raw[p] = power_law(r, deg_exp, deg_min, deg_max + 1.0);
    if (chung_lu) win[p] = power_law(r, deg_exp, deg_min, deg_max + 1.0);
  }
  // scale so that the mean out-degree is mean_deg (Chung-Lu), floor at deg_min
  double scale = 1.0;
  if (chung_lu && mean_deg > 0) {
    double lo = 1e-6, hi = 10.0;
    for (int it = 0; it < 60; ++it) {
      double mid = 0.5 * (lo + hi);
      long long s = 0;   // integer sum: independent of the thread count
#pragma omp parallel for reduction(+ : s) schedule(static)
      for (int32_t p = 0; p < n; ++p) s += (long long)std::max((double)deg_min, std::floor(raw[p] * mid));
      if ((double)s / n * (1.0 - dangling) > mean_deg) hi = mid;
      else lo = mid;
    }
    scale = 0.5 * (lo + hi);
  }
  // cumulative in-weights per community (positions) and globally
  std::vector<double> cdf;
  if (chung_lu) {
    cdf.resize(n);
    double run = 0;
    for (int32_t p = 0; p < n; ++p) {
      run += win[p];
      cdf[p] = run;
    }
  }
  // guide table of the in-weight cdf: guide[k] lies at or below the first
  // position whose cdf exceeds k * total / n, so a draw starts its search
  // there and walks a step or two instead of a binary search over all n
  // (the result is exactly std::upper_bound's)
  std::vector<int32_t> guide;
  if (chung_lu) {
    guide.resize(n);
    const double tot = cdf[n - 1];
#pragma omp parallel for schedule(static)
    for (int32_t k = 0; k < n; ++k)
      guide[k] = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), (double)k / n * tot) - cdf.begin());
  }
  auto upper = [&](double x) -> int64_t {
    const double tot = cdf[n - 1];
    int64_t k = (int64_t)(x / tot * n);
    k = k < 0 ? 0 : (k >= n ? n - 1 : k);
    int64_t t = guide[k];
    while (t > 0 && cdf[t - 1] > x) --t;
    while (t < n && cdf[t] <= x) ++t;
    return t;
  };
  Builder B(n);
  std::vector<int32_t> dpos(n);
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    Rng r(seed, (uint64_t)p, 2);
    int32_t d = 0;
    if (!(r.uniform() < dangling)) {
      double x = chung_lu ? std::max((double)deg_min, std::floor(raw[p] * scale)) : std::floor(raw[p]);
      d = (int32_t)std::min<double>(x, (double)n - 1);
      d = std::max(d, std::min(deg_min, n - 1));
    }
    dpos[p] = d;
    B.off[perm.fwd(p) + 1] = d;
  }
  B.alloc();
  const double total_w = chung_lu ? cdf[n - 1] : 0.0;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t p = 0; p < n; ++p) {
    const int32_t u = (int32_t)perm.fwd(p);
    const int32_t c = com_of_pos[p];
    const int64_t cs = cb[c], ce = cb[c + 1], csz = ce - cs;
    Rng r(seed, (uint64_t)p, 3);
    int64_t o = B.off[u];
    for (int32_t k = 0; k < dpos[p]; ++k) {
      int64_t t;
      bool inside = r.uniform() >= mu || csz == n;
      if (csz <= 1) inside = false;
      if (inside) {
        if (chung_lu) {
          double base = cs > 0 ? cdf[cs - 1] : 0.0;
          do {
            double x = base + r.uniform() * (cdf[ce - 1] - base);
            t = std::max<int64_t>(upper(x), cs);   // = upper_bound over [cs, ce): cdf[cs - 1] = base <= x
            if (t >= ce) t = ce - 1;
          } while (t == p);
        } else {
          do { t = cs + (int64_t)r.below(csz); } while (t == p);
        }
      } else {
        if (chung_lu) {
          do {
            double x = r.uniform() * total_w;
            t = upper(x);
            if (t >= n) t = n - 1;
          } while (t >= cs && t < ce);
        } else {
          t = (int64_t)r.below((uint64_t)(n - csz));
          if (t >= cs) t += csz;
        }
      }
      B.tgt[o] = (int32_t)perm.fwd((uint64_t)t);
      B.wt[o] = chung_lu ? (double)(float)std::exp(r.normal()) : 1.0;
      ++o;
    }
  }
  B.finish(*g);
  g->com.resize(n);
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) g->com[perm.fwd(p)] = com_of_pos[p];
  *out = g;
  return 0;
}

}  // namespace

extern "C" {

int rmx_syn_sbm(int32_t n, int32_t blocks, double mean_deg, double mu, uint64_t seed, rmx_syn **out) {
  if (!out || n < 2 || blocks < 1 || blocks > n || mu < 0 || mu > 1) return -22;
  rmx_syn *g = new rmx_syn();
  const int32_t size = n / blocks;
  Builder B(n);
  std::vector<int32_t> deg(n);
#pragma omp parallel for schedule(static)
  for (int32_t a = 0; a < n; ++a) {
    Rng r(seed, (uint64_t)a, 11);
    deg[a] = std::min(r.poisson(mean_deg), n - 1);
    B.off[a + 1] = deg[a];
  }
  B.alloc();
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t a = 0; a < n; ++a) {
    Rng r(seed, (uint64_t)a, 12);
    const int32_t blk = std::min(a / size, blocks - 1);
    const int32_t bs = blk * size, be = blk == blocks - 1 ? n : bs + size;
    std::vector<int32_t> t;
    t.reserve(deg[a]);
    int guard = 0;
    while ((int32_t)t.size() < deg[a] && guard < 100 * (deg[a] + 10)) {
      ++guard;
      int32_t b;
      if (r.uniform() < 1.0 - mu || be - bs == n) b = bs + (int32_t)r.below(be - bs);
      else {
        b = (int32_t)r.below(n - (be - bs));
        if (b >= bs) b += be - bs;
      }
      if (b == a || std::find(t.begin(), t.end(), b) != t.end()) continue;
      t.push_back(b);
    }
    int64_t o = B.off[a];
    for (int32_t b : t) {
      B.tgt[o] = b;
      B.wt[o] = 1.0;
      ++o;
    }
    for (; o < B.off[a + 1]; ++o) {   // unfilled slots become self-loops (dropped)
      B.tgt[o] = a;
      B.wt[o] = 1.0;
    }
  }
  B.finish(*g);
  g->com.resize(n);
  for (int32_t a = 0; a < n; ++a) g->com[a] = std::min(a / size, blocks - 1);
  *out = g;
  return 0;
}

int rmx_syn_lfr(int32_t n, double deg_exp, int32_t deg_min, int32_t deg_max, double com_exp, int32_t com_min,
                int32_t com_max, double mu, double dangling, uint64_t seed, rmx_syn **out) {
  if (!out) return -22;
  return community_model(n, deg_exp, deg_min, deg_max, 0.0, com_exp, com_min, com_max, mu, dangling, seed, false,
                         out);
}

int rmx_syn_chunglu(int32_t n, double deg_exp, int32_t deg_min, int32_t deg_max, double mean_deg, double com_exp,
                    int32_t com_min, int32_t com_max, double mu, double dangling, uint64_t seed, rmx_syn **out) {
  if (!out) return -22;
  return community_model(n, deg_exp, deg_min, deg_max, mean_deg, com_exp, com_min, com_max, mu, dangling, seed, true,
                         out);
}

int rmx_syn_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed, rmx_syn **out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1) return -22;
  const int32_t n = 1 << scale;
  const int64_t m = (int64_t)edge_factor * n;
  std::vector<int32_t> src(m), dst(m);
  std::vector<float> wv(m);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) {
    Rng r(seed, (uint64_t)e, 21);
    uint32_t s = 0, d = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      double x = r.uniform();
      uint32_t sb = 0, db = 0;
      if (x < a) { sb = 0; db = 0; }
      else if (x < a + b) { sb = 0; db = 1; }
      else if (x < a + b + c) { sb = 1; db = 0; }
      else { sb = 1; db = 1; }
      s = (s << 1) | sb;
      d = (d << 1) | db;
    }
    src[e] = (int32_t)s;
    dst[e] = (int32_t)d;
    float f = (float)r.uniform_pos();
    wv[e] = f > 0.0f ? f : 1.0f;
  }
  // bucket arcs by source in arc order (counting sort)
  Builder B(n);
  for (int64_t e = 0; e < m; ++e) B.off[src[e] + 1]++;
  B.alloc();
  {
    std::vector<int64_t> fill(B.off.begin(), B.off.end() - 1);
    for (int64_t e = 0; e < m; ++e) {
      int64_t k = fill[src[e]]++;
      B.tgt[k] = dst[e];
      B.wt[k] = (double)wv[e];
    }
  }
  std::vector<int32_t>().swap(src);
  std::vector<int32_t>().swap(dst);
  std::vector<float>().swap(wv);
  rmx_syn *g = new rmx_syn();
  B.finish(*g);
  g->com.assign(n, -1);
  *out = g;
  return 0;
}

int rmx_syn_union(rmx_syn *const *parts, int32_t k, rmx_syn **out) {
  if (!out || !parts || k < 1) return -22;
  rmx_syn *g = new rmx_syn();
  int64_t n = 0, m = 0;
  for (int32_t i = 0; i < k; ++i) {
    n += parts[i]->n;
    m += (int64_t)parts[i]->idx.size();
  }
  if (n > 0x7fffffff) {
    delete g;
    return -22;
  }
  g->n = (int32_t)n;
  g->ptr.reserve(n + 1);
  g->ptr.push_back(0);
  g->idx.reserve(m);
  g->w.reserve(m);
  g->com.reserve(n);
  int32_t noff = 0, coff = 0;
  for (int32_t i = 0; i < k; ++i) {
    const rmx_syn &p = *parts[i];
    int32_t cmax = -1;
    for (int32_t u = 0; u < p.n; ++u) {
      for (int64_t e = p.ptr[u]; e < p.ptr[u + 1]; ++e) {
        g->idx.push_back(p.idx[e] + noff);
        g->w.push_back(p.w[e]);
      }
      g->ptr.push_back((int64_t)g->idx.size());
      g->com.push_back(p.com[u] < 0 ? -1 : p.com[u] + coff);
      cmax = std::max(cmax, p.com[u]);
    }
    noff += p.n;
    coff += cmax + 1;
  }
  *out = g;
  return 0;
}

int rmx_syn_info(const rmx_syn *g, int32_t *n, int64_t *m) {
  if (!g) return -22;
  if (n) *n = g->n;
  if (m) *m = (int64_t)g->idx.size();
  return 0;
}

int rmx_syn_arrays(const rmx_syn *g, const int64_t **out_ptr, const int32_t **out_idx, const double **w,
                   const int32_t **community) {
  if (!g) return -22;
  if (out_ptr) *out_ptr = g->ptr.data();
  if (out_idx) *out_idx = g->idx.data();
  if (w) *w = g->w.data();
  if (community) *community = g->com.data();
  return 0;
}

int rmx_syn_write(const rmx_syn *g, const char *path) {
  if (!g || !path) return -22;
  std::string p(path);
  bool pajek = p.size() > 4 && p.substr(p.size() - 4) == ".net";
  FILE *f = std::fopen(path, "w");
  if (!f) return -5;
  std::vector<char> buf(1 << 22);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  if (pajek) {
    std::fprintf(f, "*Vertices %d\n", g->n);
    for (int32_t i = 0; i < g->n; ++i) std::fprintf(f, "%d \"%d\"\n", i + 1, i + 1);
    std::fprintf(f, "*Arcs %lld\n", (long long)g->idx.size());
  }
  for (int32_t u = 0; u < g->n; ++u)
    for (int64_t e = g->ptr[u]; e < g->ptr[u + 1]; ++e) {
      if (g->w[e] == 1.0) std::fprintf(f, "%d %d\n", u + 1, g->idx[e] + 1);
      else std::fprintf(f, "%d %d %.17g\n", u + 1, g->idx[e] + 1, g->w[e]);
    }
  int rc = std::fclose(f) == 0 ? 0 : -5;
  return rc;
}

void rmx_syn_free(rmx_syn *g) { delete g; }

}  // extern "C"
