// sssp driver: the reference's main() (src/sssp.cc:197-213) with DeltaStep
// replaced by gg_sssp on the GPU.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "single-source shortest-path", CLI::kDelta);
    if (!cli.ParseArgs()) return -1;
    WGraph g = MakeGraph<WNode>(cli);
    DeviceSession<WGraph> dev(cli.device(), g, cli.order_by_degree());
    int32_t delta = cli.delta();
    if (cli.autotune() && cli.num_trials() > 0) delta = Autotune(cli, dev, g, GG_TUNE_SSSP, 2).sssp_delta;
    SourcePicker<WGraph> sp(g, static_cast<NodeID>(cli.start_vertex()));
    auto SSSPBound = [&](const WGraph& gr) {
      const NodeID source = sp.PickNext();
      if (cli.logging_en()) PrintStep("Source", static_cast<int64_t>(source));
      Buf<WeightT> dist(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_sssp(dev.get(), source, delta, dist.data(), &st), "gg_sssp");
      if (cli.logging_en()) std::cout << "took " << st.iterations << " buckets" << std::endl;
      LogStats(cli.logging_en(), st);
      return dist;
    };
    SourcePicker<WGraph> vsp(g, static_cast<NodeID>(cli.start_vertex()));
    auto VerifierBound = [&](const WGraph& gr, const Buf<WeightT>& dist) {
      return SSSPVerifier(gr, vsp.PickNext(), dist.data());
    };
    auto Stats = [](const WGraph& gr, const Buf<WeightT>& dist) { PrintSSSPStats(gr, dist.data()); };
    BenchmarkKernel(cli, g, SSSPBound, Stats, VerifierBound);
    return 0;
  });
}
