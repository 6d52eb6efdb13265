// bfs driver: the reference's main() (src/bfs.cc:250-266) with DOBFS
// replaced by gg_bfs on the GPU.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "breadth-first search", CLI::kApp);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g);
    SourcePicker<Graph> sp(g, static_cast<NodeID>(cli.start_vertex()));
    auto BFSBound = [&](const Graph& gr) {
      const NodeID source = sp.PickNext();
      if (cli.logging_en()) PrintStep("Source", static_cast<int64_t>(source));
      Buf<NodeID> parent(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_bfs(dev.get(), source, 15, 18, parent.data(), nullptr, &st), "gg_bfs");
      if (cli.logging_en()) {
        PrintStep("td", st.td_steps);
        PrintStep("bu", st.bu_steps);
      }
      LogStats(cli.logging_en(), st);
      return parent;
    };
    SourcePicker<Graph> vsp(g, static_cast<NodeID>(cli.start_vertex()));
    auto VerifierBound = [&](const Graph& gr, const Buf<NodeID>& parent) {
      return BFSVerifier(gr, vsp.PickNext(), parent.data());
    };
    auto Stats = [](const Graph& gr, const Buf<NodeID>& parent) { PrintBFSStats(gr, parent.data()); };
    BenchmarkKernel(cli, g, BFSBound, Stats, VerifierBound);
    return 0;
  });
}
