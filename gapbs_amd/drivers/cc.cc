// cc driver: the reference's main() (src/cc.cc:220-229) with Afforest
// replaced by gg_cc on the GPU.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "connected-components-afforest", CLI::kApp);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    auto CCBound = [&](const Graph& gr) {
      Buf<NodeID> comp(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_cc(dev.get(), 2, comp.data(), &st), "gg_cc");
      LogStats(cli.logging_en(), st);
      return comp;
    };
    auto VerifierBound = [](const Graph& gr, const Buf<NodeID>& comp) { return CCVerifier(gr, comp.data()); };
    auto Stats = [](const Graph& gr, const Buf<NodeID>& comp) { PrintCompStats(gr, comp.data()); };
    BenchmarkKernel(cli, g, CCBound, Stats, VerifierBound);
    return 0;
  });
}
