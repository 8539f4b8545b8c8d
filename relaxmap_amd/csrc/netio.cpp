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
#include "netio.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

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

// sort by (a, b), sum duplicates in input order (stable sort keeps it)
void merge(std::vector<Arc> &arcs, int32_t &n_double) {
  std::stable_sort(arcs.begin(), arcs.end(),
                   [](const Arc &x, const Arc &y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
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
  while (ls.next(b, e)) {
    Tok t(b, e);
    t.next(buf);
    int64_t a = std::atoi(buf.c_str());
    t.next(buf);
    int64_t c = std::atoi(buf.c_str());
    buf.clear();
    t.next(buf);
    double w = buf.empty() ? 1.0 : std::atof(buf.c_str());
    arcs.push_back({a - 1, c - 1, w});
  }
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
  Lines ls{data.data(), data.data() + data.size()};
  const char *b, *e;
  std::string buf;
  std::vector<Arc> arcs;
  while (ls.next(b, e)) {
    if (b < e && *b == '#') continue;
    Tok t(b, e);
    t.next(buf);
    int64_t a = std::atol(buf.c_str());
    t.next(buf);
    int64_t c = std::atol(buf.c_str());
    buf.clear();
    t.next(buf);
    double w = buf.empty() ? 1.0 : std::atof(buf.c_str());
    arcs.push_back({a, c, w});
  }
  std::vector<int64_t> ids;
  ids.reserve(arcs.size() * 2);
  for (const Arc &x : arcs) {
    ids.push_back(x.a);
    ids.push_back(x.b);
  }
  std::sort(ids.begin(), ids.end());
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
  for (Arc &x : arcs) {
    x.a = std::lower_bound(ids.begin(), ids.end(), x.a) - ids.begin();
    x.b = std::lower_bound(ids.begin(), ids.end(), x.b) - ids.begin();
  }
  merge(arcs, net.n_double);
  build_csr(net, arcs);
  return 0;
}

}  // namespace

int read_network(const std::string &path, NetFile &net, std::string &err) {
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
