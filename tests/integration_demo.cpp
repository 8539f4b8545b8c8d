// A driver written against the reference's Network interface, run on the
// GpuNetwork binding that INTEGRATION.md shows (its code block is extracted
// to gpu_network.h at build time): the loops of stochastic_greedy_partition
// (OmpRelaxmap.cpp:263-429) for one fine phase on the nodes and one phase on
// the super nodes, with the reference's stop rule.  Prints the codelengths
// (nats, 17 digits) and module counts for the test that runs it.
#include <cmath>
#include <cstdio>
#include <exception>

#include "gpu_network.h"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: integration_demo <network file>\n");
    return 2;
  }
  try {
    GpuNetwork net(0, false);
    net.load(argv[1]);
    net.initiate(1);
    std::printf("initial %.17g %d %d\n", net.CodeLength(), net.NModule(), net.NNode());
    double tSeq = 0.0;
    const double threshold = 1e-3;
    net.resetActive();
    for (int it = 0; it < 10; ++it) {
      const double old = net.CodeLength();
      net.prioritize_parallelMove(1, tSeq, 0.0);
      if ((old - net.CodeLength()) / std::log(2.0) < threshold) break;
    }
    net.updateMembersInModule();
    std::printf("fine %.17g %d\n", net.CodeLength(), net.NModule());
    net.convertModulesToSuperNodes(1);
    net.resetActive();
    for (int it = 0; it < 10; ++it) {
      const double old = net.CodeLength();
      net.prioritize_parallelMoveSPnodes(1, tSeq, 0.0);
      if ((old - net.CodeLength()) / std::log(2.0) < threshold) break;
    }
    net.updateMembersInModule();
    std::printf("final %.17g %.17g %d\n", net.CodeLength(), net.calculateCodeLength(), net.NModule());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
