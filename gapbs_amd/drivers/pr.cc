// pr driver: the reference's main() (src/pr.cc:97-111 / src/pr_spmv.cc:97-111)
// with the PageRank kernel replaced by gg_pagerank (Jacobi pull, the
// pr_spmv.cc semantics) on the GPU.  -p fp64 selects FP64 accumulation,
// -p gs the in-place Gauss-Seidel sweep of pr.cc's PageRankPullGS.
#include "driver_common.h"

using namespace gapdrv;

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "pagerank", CLI::kPageRank, 20, 1e-4);
    if (!cli.ParseArgs()) return -1;
    Graph g = MakeGraph<NodeID>(cli);
    DeviceSession<Graph> dev(cli.device(), g, cli.order_by_degree());
    const int mode = cli.precision() == "fp64" ? GG_PR_FP64 : cli.precision() == "gs" ? GG_PR_GS : GG_PR_FP32;
    if (cli.autotune() && cli.num_trials() > 0) Autotune(cli, dev, g, GG_TUNE_PR, 0);
    auto PRBound = [&](const Graph& gr) {
      Buf<float> scores(static_cast<size_t>(gr.num_nodes()));
      gg_stats st;
      Check(gg_pagerank(dev.get(), cli.max_iters(), cli.tolerance(), mode, scores.data(), &st),
            "gg_pagerank");
      if (cli.logging_en()) {
        PrintStep("iterations", st.iterations);
        PrintTime("residual", st.residual);
      }
      LogStats(cli.logging_en(), st);
      return scores;
    };
    auto VerifierBound = [&](const Graph& gr, const Buf<float>& scores) {
      return PRVerifier(gr, scores.data(), cli.tolerance());
    };
    auto Stats = [](const Graph& gr, const Buf<float>& scores) { PrintTopScores(gr, scores.data()); };
    BenchmarkKernel(cli, g, PRBound, Stats, VerifierBound);
    return 0;
  });
}
