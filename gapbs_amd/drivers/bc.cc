// bc driver: the reference's main() (src/bc.cc:233-252) with Brandes
// replaced by gg_bc on the GPU.  The roots are drawn from the driver's
// SourcePicker, -i per trial, as Brandes draws them (bc.cc:109).
// -p fp32 (default): the reference's float arithmetic; -p kahan: compensated
// float sums; -p fp64: double dependencies and scores.  -j (> 1) runs each
// trial's -i roots (at most 64) as one batched traversal (gg_bc_batch).
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "betweenness-centrality", CLI::kIter, 1);
    if (!cli.ParseArgs()) return -1;
    if (cli.num_iters() > 1 && cli.start_vertex() != -1)
      std::cout << "Warning: iterating from same source (-r & -i)" << std::endl;
    if (cli.batch() > 1 && (cli.num_iters() > 64 || cli.precision() == "kahan")) {
      std::cout << "-j batches up to 64 roots (-i) in fp32 or fp64" << std::endl;
      return -1;
    }
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    const int mode = cli.precision() == "fp64" ? GG_BC_FP64
                     : cli.precision() == "kahan" ? GG_BC_FP32_KAHAN
                                                  : GG_BC_FP32;
    SourcePicker<Graph> sp(g, static_cast<NodeID>(cli.start_vertex()));
    auto BCBound = [&](const Graph& gr) {
      std::vector<NodeID> roots;
      for (int i = 0; i < cli.num_iters(); i++) {
        roots.push_back(sp.PickNext());
        if (cli.logging_en()) PrintStep("Source", static_cast<int64_t>(roots.back()));
      }
      Buf<float> scores(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      if (cli.batch() > 1)  // all of this trial's roots in one batched traversal
        Check(gg_bc_batch(dev.get(), roots.data(), static_cast<int>(roots.size()), mode, scores.data(), &st),
              "gg_bc_batch");
      else
        Check(gg_bc(dev.get(), roots.data(), static_cast<int>(roots.size()), mode, scores.data(), &st), "gg_bc");
      LogStats(cli.logging_en(), st);
      return scores;
    };
    SourcePicker<Graph> vsp(g, static_cast<NodeID>(cli.start_vertex()));
    auto VerifierBound = [&](const Graph& gr, const Buf<float>& scores) {
      return BCVerifier(gr, vsp, cli.num_iters(), scores.data());
    };
    auto Stats = [](const Graph& gr, const Buf<float>& scores) { PrintTopScores(gr, scores.data()); };
    BenchmarkKernel(cli, g, BCBound, Stats, VerifierBound);
    return 0;
  });
}
