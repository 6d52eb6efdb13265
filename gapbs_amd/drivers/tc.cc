// tc driver: the reference's main() (src/tc.cc:131-143) with OrderedCount
// replaced by gg_tc on the GPU.  As in Hybrid (tc.cc:95-100), a graph worth
// relabelling is relabelled by degree inside the trial; the relabelled graph
// is then uploaded and counted (also inside the trial).
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
    DeviceSession<Graph> dev(cli.device(), g);
    const bool relabel = WorthRelabelling(g);
    Graph relabelled;
    std::unique_ptr<DeviceSession<Graph>> rdev;
    auto Hybrid = [&](const Graph&) {
      uint64_t count = 0;
      gg_stats st;
      if (relabel) {
        Timer t;
        t.Start();
        relabelled = RelabelByDegree(g);
        t.Stop();
        PrintTime("Relabel", t.Seconds());
        rdev.reset(new DeviceSession<Graph>(cli.device(), relabelled));
        Check(gg_tc(rdev->get(), &count, &st), "gg_tc");
        rdev.reset();
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
