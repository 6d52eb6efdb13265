// tc driver: the reference's main() (src/tc.cc:131-143) with OrderedCount
// replaced by gg_tc on the GPU.  As in Hybrid (tc.cc:95-100), a graph worth
// relabelling is relabelled by degree inside the trial -- on the device
// (gg_relabel_by_degree) -- and the relabelled graph is counted.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "triangle count", CLI::kApp);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    if (g.directed()) {
      std::cout << "Input graph is directed but tc requires undirected" << std::endl;
      return -2;
    }
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    const bool relabel = WorthRelabelling(g);
    auto Hybrid = [&](const Graph&) {
      uint64_t count = 0;
      gg_stats st;
      if (relabel) {
        gg_graph* r = nullptr;
        gg_stats rs;
        Check(gg_relabel_by_degree(dev.get(), &r, &rs), "gg_relabel_by_degree");
        PrintTime("Relabel", rs.total_ms / 1e3);
        const int rc = gg_tc(r, &count, &st);
        gg_graph_free(r);
        Check(rc, "gg_tc");
      } else {
        Check(gg_tc(dev.get(), &count, &st), "gg_tc");
      }
      LogStats(cli.logging_en(), st);
      return count;
    };
    auto VerifierBound = [](const Graph& gr, uint64_t count) { return TCVerifier(gr, count); };
    BenchmarkKernel(cli, g, Hybrid, PrintTriangleStats, VerifierBound);
    return 0;
  });
}
