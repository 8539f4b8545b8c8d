// relaxmap: the ompRelaxmap command line (OmpRelaxmap.cpp:41-252) on the GPU.
//
//   relaxmap <seed> <network.net|.txt|.rmxcsr> <# threads> <# attempts> <threshold>
//            <vThresh> <maxIter> <outDir> <prior/normal> [selflinks]
//   relaxmap --convert <network.net|.txt> <out.rmxcsr>
//
// <# threads> is accepted and ignored (the sweeps run on one GPU);
// <# attempts> > 1 runs that many independent partitions (seeds seed,
// seed + 1, ...) and keeps the one with the lowest codelength (the reference
// parses it but makes one attempt, OmpRelaxmap.cpp:50).
// Outputs <outDir>/<name>.tree and <outDir>/<name>.clu.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "relaxmap.h"

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int die(rmx_ctx *ctx, const char *what) {
  std::printf("%s: %s\n", what, ctx ? rmx_last_error(ctx) : "no GPU context");
  if (ctx) rmx_destroy(ctx);
  return 1;
}

int main(int argc, char **argv) {
  if (argc == 4 && std::string(argv[1]) == "--convert") {
    // relaxmap --convert <network.net|.txt> <out.rmxcsr>: parse once, keep the binary CSR
    char err[512] = {0};
    rmx_net *net = nullptr;
    double t = now_s();
    if (rmx_net_read(argv[2], 0, &net, err, sizeof(err)) != RMX_OK) {
      std::printf("cannot read %s: %s\n", argv[2], err);
      return 1;
    }
    int32_t n = 0, ns = 0, nd = 0;
    int64_t m = 0;
    rmx_net_info(net, &n, &m, &ns, &nd);
    std::printf("parsed %s: %d nodes, %lld arcs in %.2f s\n", argv[2], n, (long long)m, now_s() - t);
    const int rc = rmx_net_write(net, argv[3]);
    rmx_net_free(net);
    if (rc != RMX_OK) {
      std::printf("cannot write %s\n", argv[3]);
      return 1;
    }
    return 0;
  }
  if (argc < 10) {
    std::printf("Call: ./relaxmap <seed> <network.net> <# threads> <# attempts> <threshold> <vThresh> <maxIter> "
                "<outDir> <prior/normal>  [selflinks]\n");
    return 255;
  }
  const unsigned long seed = std::strtoul(argv[1], nullptr, 10);
  const std::string file = argv[2];
  const int attempts = std::atoi(argv[4]);
  const double threshold = std::atof(argv[5]);
  const double vthresh = std::atof(argv[6]);
  const int max_iter = std::atoi(argv[7]);
  const std::string out_dir = argv[8];
  const bool prior = std::string(argv[9]) == "prior";
  const bool self_links = argc == 11 && std::string(argv[10]) == "selflinks";
  size_t slash = file.find_last_of('/');
  size_t dot = file.find_last_of('.');
  if (slash == std::string::npos) slash = 0;
  else slash += 1;
  if (dot == std::string::npos || dot < slash) dot = file.size();
  const std::string name = file.substr(slash, dot - slash);

  std::printf("Threshold = %g, Vertex-Threshold = %g\n", threshold, vthresh);
  rmx_ctx *ctx = nullptr;
  int rc = rmx_create(&ctx, 0);
  if (rc != RMX_OK) {
    std::printf("no usable GPU (status %d)\n", rc);
    return 1;
  }
  const char *prec = std::getenv("RELAXMAP_PRECISION");
  if (prec && std::string(prec) == "mixed") rmx_set_precision(ctx, RMX_MIXED);
  const char *bat = std::getenv("RELAXMAP_BATCHES");
  if (bat) rmx_set_batches(ctx, std::atoi(bat));

  double t_all = now_s();
  std::printf("Reading network %s file\n", file.c_str());
  int32_t n = 0, n_self = 0, n_double = 0;
  int64_t m = 0;
  if (rmx_load_network(ctx, file.c_str(), &n, &m, &n_self, &n_double) != RMX_OK) return die(ctx, "load");
  std::printf("done! (found %d nodes and %lld edges.)", n, (long long)m);
  if (n_double > 0) std::printf(", aggregated %d link(s) defined more than once\n", n_double);
  else std::printf("\n");
  if (self_links) std::printf("current version always excludes self links.\n");
  std::printf("ignoring %d self link(s).\n", n_self);
  double t_noio = now_s();

  int32_t iters = 0;
  double L0 = 0;
  int64_t n_dang = 0;
  double t = now_s();
  if (rmx_initiate(ctx, &iters, &L0, &n_dang) != RMX_OK) return die(ctx, "initiate");
  std::printf("Level 0: the number of dangling nodes = %lld\n", (long long)n_dang);
  std::printf("Calculating flow done in %d iterations!\n", iters);
  std::printf("DONE: Initiate() ... in %g (sec)\n", now_s() - t);
  std::printf("Initial Code Length: %g in %d modules.\n", L0 / std::log(2.0), n);
  std::printf("Now partition the network starts...\n");

  rmx_partition_opts o;
  o.threshold = threshold;
  o.vthresh = vthresh;
  o.max_iter = max_iter;
  o.prior = prior ? 1 : 0;
  o.seed = seed;
  o.verbose = 1;
  o.attempts = attempts > 1 ? attempts : 1;
  rmx_partition_result r;
  if (rmx_partition(ctx, &o, &r) != RMX_OK) return die(ctx, "partition");
  double t_end = now_s();
  std::printf("Overall Elapsed Time for Module Detection (w/o file IO): %g (sec)\n", t_end - t_noio);
  std::printf("Overall Elapsed Time for Module Detection (w/ file Reading): %g (sec)\n", t_end - t_all);
  std::printf("\nComputed Code Length = %g\n", r.computed_codelength / std::log(2.0));

  const std::string base = out_dir + "/" + name;
  if (rmx_write_tree(ctx, (base + ".tree").c_str()) != RMX_OK) return die(ctx, "tree");
  if (rmx_write_clu(ctx, (base + ".clu").c_str()) != RMX_OK) return die(ctx, "clu");
  rmx_destroy(ctx);
  return 0;
}
