// cc_sv driver: the reference's main() (src/cc_sv.cc:152-160) with
// ShiloachVishkin replaced by gg_cc_sv on the GPU.
#include <iostream>

#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "connected-components", CLI::kApp);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    auto CCBound = [&](const Graph& gr) {
      Buf<NodeID> comp(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_cc_sv(dev.get(), comp.data(), &st), "gg_cc_sv");
      std::cout << "Shiloach-Vishkin took " << st.iterations << " iterations" << std::endl;
      LogStats(cli.logging_en(), st);
      return comp;
    };
    auto VerifierBound = [](const Graph& gr, const Buf<NodeID>& comp) { return CCVerifier(gr, comp.data()); };
    auto Stats = [](const Graph& gr, const Buf<NodeID>& comp) { PrintCompStats(gr, comp.data()); };
    BenchmarkKernel(cli, g, CCBound, Stats, VerifierBound);
    return 0;
  });
}
