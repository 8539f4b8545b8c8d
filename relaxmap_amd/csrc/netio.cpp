// Network files with the reference's parsing rules.
//
// .net (FileIO.cpp:32-171): "*Vertices N"; N vertex lines whose name is the
// quoted text or the second token and whose optional weight (<= 0 -> 1)
// follows; "*Arcs" or "*Edges" (both directed); arcs "from to [w]", 1-based,
// w defaulting to 1, duplicates summed in file order.
// .txt (FileIO.cpp:177-265): '#' lines skipped; "from to [w]"; the node set is
// the union of endpoints renumbered in ascending id order, named by the id.
// Tokens follow `istringstream >> string`: a failed extraction keeps the
// previous token (FileIO.cpp:143-153, 202-214).  Then, as main() does
// (OmpRelaxmap.cpp:117-135), self-loops and w <= 0 are dropped.
// .tree (OmpRelaxmap.cpp:790-812) and .clu (OmpRelaxmap.cpp:240-249) follow
// the reference byte for byte (6 significant digits, CRLF in .clu).
//
// The arc lines are parsed by several threads, each on a run of whole lines.
// Every line clears the token buffer before its weight, so the buffer a line
// starts with is the previous arc line's third token (or empty): a thread
// finds it by looking one arc line back, and the result equals a sequential
// parse.  Duplicates are summed in file order after a stable parallel merge
// sort by (from, to).  .rmxcsr files hold the parsed CSR in binary.
#include "netio.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

namespace rmx {

namespace {

struct Arc {
  int64_t a, b;
  double w;
};

struct Tok {
  const char *p, *end;
  bool failed = false;
  Tok(const char *b, const char *e) : p(b), end(e) {}
  bool next(std::string &buf) {
    if (failed) return false;
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    if (p >= end) {
      failed = true;
      return false;
    }
    const char *s = p;
    while (p < end && !std::isspace((unsigned char)*p)) ++p;
    buf.assign(s, p);
    return true;
  }
};

struct Lines {
  const char *p, *end;
  bool next(const char *&b, const char *&e) {
    if (p >= end) return false;
    b = p;
    const char *q = (const char *)std::memchr(p, '\n', (size_t)(end - p));
    if (q) {
      e = q;
      p = q + 1;
    } else {
      e = end;
      p = end;
    }
    return true;
  }
};

bool slurp(const std::string &path, std::string &data) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  data = ss.str();
  return true;
}

thread_local unsigned g_threads = 0;   // 0: hardware concurrency (per calling thread)

unsigned n_threads(size_t work) {
  unsigned t = g_threads ? g_threads : std::thread::hardware_concurrency();
  if (t < 1) t = 1;
  if (t > 64) t = 64;
  const size_t most = work / 4096 + 1;   // no thread for less than a few thousand items
  return (unsigned)std::min<size_t>(t, most);
}

template <typename F>
void run_threads(unsigned t, F &&f) {
  std::vector<std::thread> th;
  for (unsigned k = 1; k < t; ++k) th.emplace_back(f, k);
  f(0u);
  for (auto &x : th) x.join();
}

bool arc_less(const Arc &x, const Arc &y) { return x.a != y.a ? x.a < y.a : x.b < y.b; }

// stable: runs sorted in parallel, then merged pairwise (std::merge takes
// equal elements from the first run first)
template <typename T, typename Less>
void parallel_stable_sort(std::vector<T> &v, Less less) {
  const unsigned t = n_threads(v.size());
  if (t <= 1) {
    std::stable_sort(v.begin(), v.end(), less);
    return;
  }
  std::vector<size_t> bnd(t + 1);
  for (unsigned k = 0; k <= t; ++k) bnd[k] = v.size() * k / t;
  run_threads(t, [&](unsigned k) { std::stable_sort(v.begin() + bnd[k], v.begin() + bnd[k + 1], less); });
  std::vector<T> tmp(v.size());
  std::vector<T> *src = &v, *dst = &tmp;
  while (bnd.size() > 2) {
    const size_t runs = bnd.size() - 1;
    std::vector<size_t> nb;
    for (size_t r = 0; r < runs; r += 2) nb.push_back(bnd[r]);
    nb.push_back(bnd.back());
    const unsigned pairs = (unsigned)((runs + 1) / 2);
    run_threads(pairs, [&](unsigned q) {
      const size_t r = 2 * (size_t)q;
      if (r + 1 < runs)
        std::merge(src->begin() + bnd[r], src->begin() + bnd[r + 1], src->begin() + bnd[r + 1],
                   src->begin() + bnd[r + 2], dst->begin() + bnd[r], less);
      else
        std::copy(src->begin() + bnd[r], src->begin() + bnd[r + 1], dst->begin() + bnd[r]);
    });
    std::swap(src, dst);
    bnd.swap(nb);
  }
  if (src != &v) v.swap(*src);
}

// sort by (a, b), sum duplicates in input order (the stable sort keeps it)
void merge(std::vector<Arc> &arcs, int32_t &n_double) {
  parallel_stable_sort(arcs, arc_less);
  size_t u = 0;
  n_double = 0;
  for (size_t i = 0; i < arcs.size(); ++i) {
    if (u > 0 && arcs[u - 1].a == arcs[i].a && arcs[u - 1].b == arcs[i].b) {
      double s = arcs[u - 1].w + arcs[i].w;
      if (s > arcs[i].w) ++n_double;
      arcs[u - 1].w = s;
    } else {
      double s = 0.0 + arcs[i].w;
      if (s > arcs[i].w) ++n_double;
      arcs[u] = arcs[i];
      arcs[u].w = s;
      ++u;
    }
  }
  arcs.resize(u);
}

// The token buffer after one arc line: its third token, or empty.
std::string carry_after(const char *b, const char *e) {
  Tok t(b, e);
  std::string buf;
  t.next(buf);
  t.next(buf);
  buf.clear();
  t.next(buf);
  return buf;
}

// One arc line "from to [w]" with the reference's token semantics; `buf`
// enters with the previous line's carry and leaves with this line's.
template <bool LONG>
inline Arc parse_arc(const char *b, const char *e, std::string &buf, int64_t offset) {
  Tok t(b, e);
  t.next(buf);
  const int64_t a = LONG ? (int64_t)std::atol(buf.c_str()) : (int64_t)std::atoi(buf.c_str());
  t.next(buf);
  const int64_t c = LONG ? (int64_t)std::atol(buf.c_str()) : (int64_t)std::atoi(buf.c_str());
  buf.clear();
  t.next(buf);
  const double w = buf.empty() ? 1.0 : std::atof(buf.c_str());
  return Arc{a - offset, c - offset, w};
}

// All arc lines of [b, e), in file order, parsed by several threads.
// `carry0` is the buffer entering the first line; `skip_hash` skips lines
// starting with '#' (they leave the buffer alone).
template <bool LONG>
void parse_arc_lines(const char *b, const char *e, const std::string &carry0, bool skip_hash, int64_t offset,
                     std::vector<Arc> &out) {
  const size_t bytes = (size_t)(e - b);
  const unsigned t = n_threads(bytes / 16);
  std::vector<const char *> cut(t + 1);
  cut[0] = b;
  cut[t] = e;
  for (unsigned k = 1; k < t; ++k) {
    const char *p = b + bytes * k / t;
    if (p < cut[k - 1]) p = cut[k - 1];
    const char *q = (const char *)std::memchr(p, '\n', (size_t)(e - p));
    cut[k] = q ? q + 1 : e;
  }
  std::vector<std::vector<Arc>> part(t);
  run_threads(t, [&](unsigned k) {
    const char *lo = cut[k], *hi = cut[k + 1];
    if (lo >= hi) return;
    // the carry entering this run: the previous arc line's, looking back
    std::string buf = carry0;
    const char *end_prev = lo;
    while (end_prev > b) {
      const char *le = end_prev - 1;   // the '\n' ending the previous line
      const char *ls = le;
      while (ls > b && ls[-1] != '\n') --ls;
      if (!(skip_hash && ls < le && *ls == '#')) {
        buf = carry_after(ls, le);
        break;
      }
      end_prev = ls;
    }
    std::vector<Arc> &v = part[k];
    v.reserve((size_t)(hi - lo) / 12 + 1);
    Lines ls{lo, hi};
    const char *lb, *le;
    while (ls.next(lb, le)) {
      if (skip_hash && lb < le && *lb == '#') continue;
      v.push_back(parse_arc<LONG>(lb, le, buf, offset));
    }
  });
  size_t total = out.size();
  for (auto &v : part) total += v.size();
  out.reserve(total);
  for (auto &v : part) out.insert(out.end(), v.begin(), v.end());
}

void build_csr(NetFile &net, const std::vector<Arc> &arcs) {
  net.ptr.assign((size_t)net.n + 1, 0);
  net.n_self = 0;
  for (const Arc &x : arcs) {
    if (!(x.w > 0.0)) continue;
    if (x.a == x.b) ++net.n_self;
    else ++net.ptr[x.a + 1];
  }
  for (int32_t i = 0; i < net.n; ++i) net.ptr[i + 1] += net.ptr[i];
  net.idx.resize(net.ptr[net.n]);
  net.w.resize(net.ptr[net.n]);
  size_t k = 0;
  for (const Arc &x : arcs) {
    if (!(x.w > 0.0) || x.a == x.b) continue;
    net.idx[k] = (int32_t)x.b;
    net.w[k] = x.w;
    ++k;
  }
}

bool is_vertices(const std::string &s) { return s == "*Vertices" || s == "*vertices" || s == "*VERTICES"; }
bool is_arcs(const std::string &s) { return s == "*Edges" || s == "*edges" || s == "*Arcs" || s == "*arcs"; }

int read_net(const std::string &data, NetFile &net, std::string &err) {
  Lines ls{data.data(), data.data() + data.size()};
  const char *b, *e;
  std::string buf;
  int n = 0;
  while (n == 0) {
    if (!ls.next(b, e)) {
      err = "the network file is not in Pajek format";
      return -1;
    }
    Tok t(b, e);
    t.next(buf);
    if (is_vertices(buf)) {
      t.next(buf);
      n = std::atoi(buf.c_str());
    } else {
      err = "the network file is not in Pajek format";
      return -1;
    }
  }
  if (n < 0) {
    err = "negative vertex count";
    return -1;
  }
  net.n = n;
  net.names.assign(n, std::string());
  net.node_w.assign(n, 1.0);
  for (int i = 0; i < n; ++i) {
    if (!ls.next(b, e)) b = e = data.data() + data.size();
    const char *q1 = nullptr, *q2 = nullptr;
    for (const char *c = b; c < e; ++c)
      if (*c == '"') {
        if (!q1) q1 = c;
        q2 = c;
      }
    Tok t(b, e);
    t.next(buf);   // id (ignored: vertices are taken in file order)
    if (q1 && q1 < q2) {
      net.names[i].assign(q1 + 1, q2);
      t = Tok(q2 + 1, e);
    } else {
      t = Tok(b, e);
      t.next(buf);
      t.next(buf);
      net.names[i] = buf;
    }
    buf = "1";
    t.next(buf);
    double nw = std::atof(buf.c_str());
    net.node_w[i] = nw <= 0.0 ? 1.0 : nw;
  }
  {
    if (!ls.next(b, e)) b = e = data.data() + data.size();
    Tok t(b, e);
    t.next(buf);
    if (!is_arcs(buf)) {
      err = "Number of nodes not matching";
      return -1;
    }
  }
  std::vector<Arc> arcs;
  parse_arc_lines<false>(ls.p, ls.end, buf, false, 1, arcs);
  merge(arcs, net.n_double);
  for (const Arc &x : arcs)
    if (x.w > 0.0 && x.a != x.b && (x.a < 0 || x.a >= n || x.b < 0 || x.b >= n)) {
      err = "arc " + std::to_string(x.a + 1) + " -> " + std::to_string(x.b + 1) + " outside 1.." + std::to_string(n);
      return -1;
    }
  build_csr(net, arcs);
  return 0;
}

int read_txt(const std::string &data, NetFile &net, std::string &err) {
  std::vector<Arc> arcs;
  parse_arc_lines<true>(data.data(), data.data() + data.size(), std::string(), true, 0, arcs);
  std::vector<int64_t> ids(arcs.size() * 2);
  {
    const unsigned t = n_threads(arcs.size());
    run_threads(t, [&](unsigned k) {
      for (size_t i = arcs.size() * k / t; i < arcs.size() * (k + 1) / t; ++i) {
        ids[2 * i] = arcs[i].a;
        ids[2 * i + 1] = arcs[i].b;
      }
    });
  }
  parallel_stable_sort(ids, [](int64_t x, int64_t y) { return x < y; });
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  if (ids.empty()) {
    err = "network has no nodes";
    return -1;
  }
  if (ids.size() > (size_t)0x7fffffff) {
    err = "too many nodes";
    return -1;
  }
  net.n = (int32_t)ids.size();
  net.names.resize(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) net.names[i] = std::to_string(ids[i]);
  net.node_w.assign(ids.size(), 1.0);
  {
    const unsigned t = n_threads(arcs.size());
    run_threads(t, [&](unsigned k) {
      for (size_t i = arcs.size() * k / t; i < arcs.size() * (k + 1) / t; ++i) {
        arcs[i].a = std::lower_bound(ids.begin(), ids.end(), arcs[i].a) - ids.begin();
        arcs[i].b = std::lower_bound(ids.begin(), ids.end(), arcs[i].b) - ids.begin();
      }
    });
  }
  merge(arcs, net.n_double);
  build_csr(net, arcs);
  return 0;
}

}  // namespace

void set_parse_threads(unsigned t) { g_threads = t; }

// .rmxcsr: "RMXCSR01", n, n_self, n_double, flags (i32 each), m (i64),
// out_ptr (i64, n + 1), out_idx (i32, m), w (f64, m), node_w (f64, n), then
// the names as one block of NUL-terminated strings preceded by its size.
namespace {
const char kMagic[8] = {'R', 'M', 'X', 'C', 'S', 'R', '0', '1'};

template <typename T>
bool put(std::ofstream &f, const T *p, size_t n) {
  f.write(reinterpret_cast<const char *>(p), (std::streamsize)(sizeof(T) * n));
  return f.good();
}
template <typename T>
bool get(std::ifstream &f, T *p, size_t n) {
  f.read(reinterpret_cast<char *>(p), (std::streamsize)(sizeof(T) * n));
  return f.good();
}
}  // namespace

int write_binary(const std::string &path, const NetFile &net, std::string &err) {
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot write " + path;
    return -2;
  }
  const int32_t hdr[4] = {net.n, net.n_self, net.n_double, 1};
  const int64_t m = (int64_t)net.idx.size();
  std::string blob;
  for (const std::string &x : net.names) {
    blob += x;
    blob.push_back('\0');
  }
  const int64_t nb = (int64_t)blob.size();
  bool ok = put(f, kMagic, 8) && put(f, hdr, 4) && put(f, &m, 1) && put(f, net.ptr.data(), net.ptr.size()) &&
            put(f, net.idx.data(), net.idx.size()) && put(f, net.w.data(), net.w.size()) &&
            put(f, net.node_w.data(), net.node_w.size()) && put(f, &nb, 1) && put(f, blob.data(), blob.size());
  if (!ok) {
    err = "cannot write " + path;
    return -2;
  }
  return 0;
}

int read_binary(const std::string &path, NetFile &net, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return -2;
  }
  char magic[8];
  int32_t hdr[4];
  int64_t m = 0, nb = 0;
  if (!get(f, magic, 8) || std::memcmp(magic, kMagic, 8) != 0 || !get(f, hdr, 4) || !get(f, &m, 1) || hdr[0] < 0 ||
      m < 0) {
    err = path + " is not a relaxmap binary network";
    return -1;
  }
  net.n = hdr[0];
  net.n_self = hdr[1];
  net.n_double = hdr[2];
  net.ptr.resize((size_t)net.n + 1);
  net.idx.resize((size_t)m);
  net.w.resize((size_t)m);
  net.node_w.resize((size_t)net.n);
  if (!get(f, net.ptr.data(), net.ptr.size()) || !get(f, net.idx.data(), net.idx.size()) ||
      !get(f, net.w.data(), net.w.size()) || !get(f, net.node_w.data(), net.node_w.size()) || !get(f, &nb, 1) ||
      nb < 0) {
    err = path + " is truncated";
    return -1;
  }
  std::string blob((size_t)nb, '\0');
  if (nb > 0 && !get(f, &blob[0], (size_t)nb)) {
    err = path + " is truncated";
    return -1;
  }
  net.names.clear();
  net.names.reserve((size_t)net.n);
  size_t p = 0;
  for (int32_t i = 0; i < net.n; ++i) {
    const size_t q = blob.find('\0', p);
    if (q == std::string::npos) {
      err = path + " has too few names";
      return -1;
    }
    net.names.emplace_back(blob, p, q - p);
    p = q + 1;
  }
  if (net.ptr[0] != 0 || net.ptr[(size_t)net.n] != m) {
    err = path + " has an inconsistent CSR";
    return -1;
  }
  for (int64_t k = 0; k < m; ++k)
    if (net.idx[(size_t)k] < 0 || net.idx[(size_t)k] >= net.n) {
      err = path + " has an arc outside the node range";
      return -1;
    }
  return 0;
}

int read_network(const std::string &path, NetFile &net, std::string &err) {
  {
    const size_t d = path.find_last_of('.');
    if (d != std::string::npos && path.substr(d) == ".rmxcsr") return read_binary(path, net, err);
  }
  std::string data;
  if (!slurp(path, data)) {
    err = "cannot open " + path;
    return -2;
  }
  size_t dot = path.find_last_of('.');
  bool pajek = dot != std::string::npos && path.substr(dot) == ".net";
  return pajek ? read_net(data, net, err) : read_txt(data, net, err);
}

int write_tree(const std::string &path, double codelength_nats, const std::vector<uint32_t> &mod,
               const std::vector<double> &p, const std::vector<std::string> &names, std::string &err) {
  std::ofstream f(path.c_str());
  if (!f) {
    err = "cannot write " + path;
    return -2;
  }
  const size_t n = mod.size();
  std::vector<uint32_t> cnt(n + 1, 0);
  for (uint32_t m : mod) ++cnt[m + 1];
  size_t n_module = 0;
  for (size_t m = 0; m < n; ++m) n_module += cnt[m + 1] > 0;
  for (size_t m = 0; m < n; ++m) cnt[m + 1] += cnt[m];
  std::vector<uint32_t> members(n), fill(cnt.begin(), cnt.end() - 1);
  for (size_t i = 0; i < n; ++i) members[fill[mod[i]]++] = (uint32_t)i;
  f << "# Code length " << codelength_nats / std::log(2.0) << " in " << n_module << " modules." << std::endl;
  int rank = 0;
  for (size_t m = 0; m < n; ++m) {
    uint32_t b = cnt[m], e = cnt[m + 1];
    if (b == e) continue;
    ++rank;
    for (uint32_t j = b; j < e; ++j) {
      uint32_t a = members[j];
      f << rank << ":" << (j - b + 1) << " " << p[a] << " \"" << names[a] << "\"" << '\n';
    }
  }
  return f.good() ? 0 : -2;
}

int write_clu(const std::string &path, const std::vector<uint32_t> &mod, std::string &err) {
  std::ofstream f(path.c_str(), std::ios::binary);
  if (!f) {
    err = "cannot write " + path;
    return -2;
  }
  f << "*Vertices " << mod.size() << "\x0D\x0A";
  for (uint32_t m : mod) f << (m + 1) << "\x0D\x0A";
  return f.good() ? 0 : -2;
}

}  // namespace rmx
