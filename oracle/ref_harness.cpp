// Golden-vector harness over the reference implementation.
//
// TEST INFRASTRUCTURE ONLY.  This file is linked against the reference's own
// compiled sources (Module.cpp, Node.cpp, FileIO.cpp, timing.cpp and
// OmpRelaxmap.cpp, the latter with its `main` renamed at compile time) by
// oracle/Makefile.  It drives the reference's `Network` API exactly the way
// the reference's `main` does (OmpRelaxmap.cpp:93-162) and prints the results
// with 17 significant digits as JSON, so that tests/golden/ can hold
// bit-faithful fixtures.  Nothing here is shipped or measured.
//
// Commands (all sequential: numTh == 1, which makes the reference
// deterministic for a given seed):
//   flows      <net>                      steady state + singleton modules
//   codelength <net> <part>               calculateCodeLength of a partition
//   contract   <net> <part>               convertModulesToSuperNodes
//   sweeps     <net> <seed> <k> <variant> k sweeps from singletons
//                                          variant: move | prior | pmove | ppmove
//   spsweeps   <net> <part> <seed> <k> <variant>  contract then k SP sweeps
//                                          variant: move | prior | pmove | ppmove
//   run        <net> <seed> <thr> <vthr> <maxIter> <prior|normal>
//                                          main()'s super-step loop, stopping
//                                          before the first coarse tune
//   write      <net> <part> <dir> <name>  the reference's .tree writer
//                                          (print_twoLevel_Cluster) on an
//                                          installed partition: <dir>/<name>.tree
// <part> is a text file with one 0-based module index per node.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "Module.h"
#include "Node.h"
#include "FileIO.h"

void stochastic_greedy_partition(Network &network, int numTh, double threshold,
                                 double vThresh, int maxIter, bool prior,
                                 bool fineTune, bool fast);
void print_twoLevel_Cluster(Network network, std::string networkName, std::string outDir);

static std::stringstream g_sink;
static std::streambuf *g_saved = 0;
static void mute() { g_sink.str(""); g_saved = std::cout.rdbuf(g_sink.rdbuf()); }
static std::string unmute() {
  std::cout.rdbuf(g_saved);
  std::string s = g_sink.str();
  g_sink.str("");
  return s;
}

static void pd(double x) { std::printf("%.17g", x); }

static void print_darray(const char *key, const std::vector<double> &v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) { if (i) std::printf(", "); pd(v[i]); }
  std::printf("]");
}
static void print_iarray(const char *key, const std::vector<long> &v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) { if (i) std::printf(", "); std::printf("%ld", v[i]); }
  std::printf("]");
}

// Mirrors OmpRelaxmap.cpp:93-162 (load, teleport weights, out-links with
// self-loops and w <= 0 dropped, initiate, in-links).
static int setup(Network &net, const std::string &file, unsigned long seed, std::string *log) {
  net.R = new MTRand(seed);
  std::string type = file.substr(file.find_last_of('.'));
  mute();
  if (type == ".net") load_pajek_format_network(file, net);
  else load_linkList_format_network(file, net);
  int n = net.NNode();
  double tot = net.TotNodeWeights();
  for (int i = 0; i < n; i++) net.nodes[i].setTeleportWeight(net.nodes[i].NodeWeight() / tot);
  for (std::map<std::pair<int, int>, double>::iterator it = net.Edges.begin(); it != net.Edges.end(); it++) {
    int from = it->first.first, to = it->first.second;
    double w = it->second;
    if (w > 0.0 && from != to) net.nodes[from].outLinks.push_back(std::make_pair(to, w));
  }
  std::map<std::pair<int, int>, double>().swap(net.Edges);
  net.initiate(1);
  for (int i = 0; i < n; i++) {
    int k = net.nodes[i].outLinks.size();
    for (int j = 0; j < k; j++)
      net.nodes[net.nodes[i].outLinks[j].first].inLinks.push_back(std::make_pair(i, net.nodes[i].outLinks[j].second));
  }
  std::string s = unmute();
  if (log) *log = s;
  return n;
}

static int parse_iters(const std::string &log) {
  size_t p = log.find("Calculating flow done in ");
  if (p == std::string::npos) return -1;
  return std::atoi(log.c_str() + p + 25);
}

static std::vector<long> read_part(const std::string &file, int n) {
  std::ifstream in(file.c_str());
  std::vector<long> part;
  long x;
  while (in >> x) part.push_back(x);
  if ((int)part.size() != n) {
    std::fprintf(stderr, "partition has %d entries, network has %d nodes\n", (int)part.size(), n);
    std::exit(2);
  }
  return part;
}

// Installs a partition: per-module sums in ascending node order, then the
// reference's own updateMembersInModule and updateCodeLength (the latter is
// unused by the reference CLI but sets each module's exitPr from scratch,
// Module.cpp:2918-2962).
static void install_partition(Network &net, const std::vector<long> &part) {
  int n = net.NNode();
  for (int m = 0; m < n; m++) {
    net.modules[m].setIndex(m);
    net.modules[m].setSumPr(0.0);
    net.modules[m].setSumTPWeight(0.0);
    net.modules[m].setSumDangling(0.0);
    net.modules[m].setNumMembers(0);
    net.modules[m].setExitPr(0.0);
    net.modules[m].setStayPr(0.0);
  }
  int nmod = 0;
  for (int i = 0; i < n; i++) {
    int m = (int)part[i];
    net.nodes[i].setModIdx(m);
    if (net.modules[m].NumMembers() == 0) nmod++;
    net.modules[m].increaseNumMembers();
    net.modules[m].addSumPr(net.nodes[i].Size());
    net.modules[m].addSumTPWeight(net.nodes[i].TeleportWeight());
    if (net.nodes[i].IsDangling()) net.modules[m].addSumDangling(net.nodes[i].Size());
  }
  net.emptyModules.clear();
  for (int m = 0; m < n; m++)
    if (net.modules[m].NumMembers() == 0) net.emptyModules.push_back(m);
  net.setNModule(nmod);
  net.setNEmptyMod(n - nmod);
  net.updateMembersInModule();
  net.updateCodeLength(1, false);
}

static void print_state(Network &net, bool with_part) {
  int n = net.NNode();
  std::printf("\"codelength\": "); pd(net.CodeLength());
  std::printf(", \"sum_exit\": "); pd(net.SumAllExitPr());
  std::printf(", \"n_module\": %d", net.NModule());
  std::printf(", \"n_empty\": %d", net.NEmptyMod());
  if (with_part) {
    std::vector<long> part(n);
    for (int i = 0; i < n; i++) part[i] = net.nodes[i].ModIdx();
    std::printf(", ");
    print_iarray("partition", part);
  }
}

static void print_supernodes(Network &net) {
  int s = net.superNodes.size();
  std::printf("\"n_super\": %d, \"super\": [", s);
  for (int i = 0; i < s; i++) {
    SuperNode &sp = net.superNodes[i];
    if (i) std::printf(",\n");
    std::printf("{\"size\": "); pd(sp.Size());
    std::printf(", \"exit\": "); pd(sp.ExitPr());
    std::printf(", \"tp\": "); pd(sp.TeleportWeight());
    std::printf(", \"dangling\": "); pd(sp.DanglingSize());
    std::printf(", \"mod\": %d, \"n_members\": %d, \"out\": [", sp.ModIdx(), (int)sp.members.size());
    for (size_t j = 0; j < sp.outLinks.size(); j++) {
      if (j) std::printf(", ");
      std::printf("[%d, ", sp.outLinks[j].first); pd(sp.outLinks[j].second); std::printf("]");
    }
    std::printf("], \"in\": [");
    for (size_t j = 0; j < sp.inLinks.size(); j++) {
      if (j) std::printf(", ");
      std::printf("[%d, ", sp.inLinks[j].first); pd(sp.inLinks[j].second); std::printf("]");
    }
    std::printf("]}");
  }
  std::printf("]");
}

static int sweep_once(Network &net, const std::string &variant, bool sp) {
  double t = 0.0;
  if (!sp) {
    if (variant == "move") return net.move();
    if (variant == "prior") return net.prioritize_move(0.0);
    if (variant == "pmove") return net.parallelMove(1, t);
    if (variant == "ppmove") return net.prioritize_parallelMove(1, t, 0.0);
  } else {
    if (variant == "move") return net.moveSuperNodes();
    if (variant == "prior") return net.prioritize_moveSPnodes(0.0);
    if (variant == "pmove") return net.parallelMoveSuperNodes(1, t);
    if (variant == "ppmove") return net.prioritize_parallelMoveSPnodes(1, t, 0.0);
  }
  std::fprintf(stderr, "unknown variant %s\n", variant.c_str());
  std::exit(2);
}

static void reset_active(Network &net, int n) {
  std::vector<char>(n).swap(net.isActives);
  std::vector<int>(n).swap(net.activeNodes);
  for (int i = 0; i < n; i++) { net.activeNodes[i] = i; net.isActives[i] = 0; }
}

int main(int argc, char **argv) {
  if (argc < 3) { std::fprintf(stderr, "usage: see header\n"); return 2; }
  std::string cmd = argv[1], file = argv[2];
  Network net;
  if (cmd == "flows") {
    std::string log;
    int n = setup(net, file, 1, &log);
    std::vector<double> p(n), q(n);
    for (int i = 0; i < n; i++) { p[i] = net.nodes[i].Size(); q[i] = net.modules[i].ExitPr(); }
    std::printf("{\"n\": %d, \"iterations\": %d, \"n_dangling\": %d, ", n, parse_iters(log), net.NDanglings());
    std::printf("\"all_log_all\": "); pd(net.AllLogAll()); std::printf(", ");
    print_state(net, false); std::printf(", ");
    print_darray("p", p); std::printf(", ");
    print_darray("q", q); std::printf("}\n");
  } else if (cmd == "codelength" && argc >= 4) {
    int n = setup(net, file, 1, 0);
    install_partition(net, read_part(argv[3], n));
    std::printf("{\"calculated\": "); pd(net.calculateCodeLength());
    std::printf(", ");
    print_state(net, false);
    std::printf("}\n");
  } else if (cmd == "contract" && argc >= 4) {
    int n = setup(net, file, 1, 0);
    install_partition(net, read_part(argv[3], n));
    mute();
    net.convertModulesToSuperNodes(1);
    unmute();
    std::printf("{");
    print_supernodes(net);
    std::printf("}\n");
  } else if (cmd == "sweeps" && argc >= 6) {
    int n = setup(net, file, std::strtoul(argv[3], 0, 10), 0);
    int k = std::atoi(argv[4]);
    std::string variant = argv[5];
    reset_active(net, n);
    std::printf("{\"init_codelength\": "); pd(net.CodeLength());
    std::printf(", \"sweeps\": [");
    for (int s = 0; s < k; s++) {
      mute();
      int moved = sweep_once(net, variant, false);
      unmute();
      if (s) std::printf(",\n");
      std::printf("{\"moved\": %d, \"n_active_next\": %d, ", moved, (int)net.activeNodes.size());
      print_state(net, true);
      std::printf("}");
    }
    std::printf("]}\n");
  } else if (cmd == "spsweeps" && argc >= 7) {
    int n = setup(net, file, std::strtoul(argv[4], 0, 10), 0);
    install_partition(net, read_part(argv[3], n));
    int k = std::atoi(argv[5]);
    std::string variant = argv[6];
    mute();
    net.convertModulesToSuperNodes(1);
    unmute();
    reset_active(net, net.superNodes.size());
    std::printf("{\"init_codelength\": "); pd(net.CodeLength());
    std::printf(", \"n_super\": %d, \"sweeps\": [", (int)net.superNodes.size());
    for (int s = 0; s < k; s++) {
      mute();
      int moved = sweep_once(net, variant, true);
      unmute();
      if (s) std::printf(",\n");
      std::printf("{\"moved\": %d, ", moved);
      print_state(net, true);
      std::printf("}");
    }
    std::printf("]}\n");
  } else if (cmd == "run" && argc >= 8) {
    // OmpRelaxmap.cpp:181-210 with numTh == 1, ending where the first coarse
    // tune (seeded from /dev/urandom, OmpRelaxmap.cpp:529) would start.
    int n = setup(net, file, std::strtoul(argv[3], 0, 10), 0);
    double threshold = std::atof(argv[4]);
    double vThresh = -std::atof(argv[5]);
    int maxIter = std::atoi(argv[6]);
    bool prior = std::string(argv[7]) == "prior";
    std::printf("{\"n\": %d, \"init_codelength\": ", n); pd(net.CodeLength());
    std::printf(", \"steps\": [");
    bool fineTune = true, nextIter = true;
    double old = net.CodeLength();
    int step = 0;
    bool coarse_next = false;
    while (nextIter) {
      old = net.CodeLength();
      mute();
      stochastic_greedy_partition(net, 1, threshold, vThresh, maxIter, prior, fineTune, false);
      std::string log = unmute();
      if (step) std::printf(",\n");
      std::printf("{\"codelength\": "); pd(net.CodeLength());
      std::printf(", \"n_module\": %d, \"moved\": [", net.NModule());
      // per-sweep numMoved values as printed by stochastic_greedy_partition
      size_t pos = 0; bool first = true;
      while ((pos = log.find("numMoved:\t", pos)) != std::string::npos) {
        pos += 10;
        std::printf(first ? "%d" : ", %d", std::atoi(log.c_str() + pos));
        first = false;
      }
      std::printf("]}");
      step++;
      if ((old - net.CodeLength()) / std::log(2.0) < threshold) nextIter = false;
      if (step >= 2) {
        fineTune = !fineTune;
        if (nextIter && !fineTune) { coarse_next = true; break; }
      }
    }
    std::printf("], \"stopped_before_coarse\": %s, \"calculated\": ", coarse_next ? "true" : "false");
    pd(net.calculateCodeLength());
    std::vector<long> part(n);
    for (int i = 0; i < n; i++) part[i] = net.nodes[i].ModIdx();
    std::printf(", ");
    print_iarray("partition", part);
    std::printf("}\n");
  } else if (cmd == "write" && argc >= 6) {
    // OmpRelaxmap.cpp:785-808 on an installed partition (its codelength and
    // module exits from updateCodeLength, its member lists ascending)
    int n = setup(net, file, 1, 0);
    install_partition(net, read_part(argv[3], n));
    print_twoLevel_Cluster(net, argv[5], argv[4]);
    std::printf("{\"codelength\": "); pd(net.CodeLength());
    std::printf(", \"n_module\": %d}\n", net.NModule());
  } else {
    std::fprintf(stderr, "bad command\n");
    return 2;
  }
  return 0;
}
