// bfs driver: the reference's main() (src/bfs.cc:250-266) with DOBFS
// replaced by gg_bfs on the GPU; -j k runs k roots per trial in one batched
// traversal.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "breadth-first search", CLI::kApp);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    int alpha = 15, beta = 18;  // DOBFS's defaults (src/bfs.cc:122-123)
    if (cli.autotune() && cli.num_trials() > 0) {
      if (cli.batch() > 1) {
        alpha = Autotune(cli, dev, g, GG_TUNE_BATCH, std::min(cli.batch(), 64)).batch_alpha;
      } else {
        const gg_tuning t = Autotune(cli, dev, g, GG_TUNE_BFS, 4);
        alpha = t.bfs_alpha;
        beta = t.bfs_beta;
      }
    }
    SourcePicker<Graph> sp(g, static_cast<NodeID>(cli.start_vertex()));
    auto BFSBound = [&](const Graph& gr) {
      const NodeID source = sp.PickNext();
      if (cli.logging_en()) PrintStep("Source", static_cast<int64_t>(source));
      Buf<NodeID> parent(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_bfs(dev.get(), source, alpha, beta, parent.data(), nullptr, &st), "gg_bfs");
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
    if (cli.batch() <= 1) {
      BenchmarkKernel(cli, g, BFSBound, Stats, VerifierBound);
      return 0;
    }
    // -j k: each trial runs the next k roots as one batched traversal
    // (gg_bfs_batch); the result is their depths, verified root by root
    // against the serial BFS of the reference's verifier.
    const int k = std::min(cli.batch(), 64);
    auto BatchBound = [&](const Graph& gr) {
      std::vector<int32_t> roots;
      for (int i = 0; i < k; i++) roots.push_back(sp.PickNext());
      Buf<int32_t> depths(static_cast<size_t>(gr.num_nodes()) * k);
      gg_stats st;
      Check(gg_bfs_batch(dev.get(), roots.data(), k, alpha, depths.data(), &st), "gg_bfs_batch");
      if (cli.logging_en()) {
        PrintStep("roots", static_cast<int64_t>(k));
        PrintStep("td", st.td_steps);
        PrintStep("bu", st.bu_steps);
      }
      LogStats(cli.logging_en(), st);
      return depths;
    };
    auto BatchVerifier = [&](const Graph& gr, const Buf<int32_t>& depths) {
      bool ok = true;
      for (int i = 0; i < k; i++) {
        const std::vector<int32_t> want = SerialBFSDepths(gr, vsp.PickNext());
        if (!std::equal(want.begin(), want.end(), depths.data() + static_cast<size_t>(i) * gr.num_nodes()))
          ok = false;
      }
      return ok;
    };
    auto BatchStats = [&](const Graph& gr, const Buf<int32_t>& depths) {
      int64_t reached = 0;
      for (size_t i = 0; i < depths.size(); i++) reached += depths[i] >= 0;
      PrintStep("reached", reached);
    };
    BenchmarkKernel(cli, g, BatchBound, BatchStats, BatchVerifier);
    return 0;
  });
}
