// Test infrastructure only: the reference's PageRankPullGS (src/pr.cc:35-61)
// behind a C-ABI for parity tests and the CPU baseline.
#define main gapbs_reference_main
#include "pr.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                    const int64_t* ioff, const int32_t* inbr) {
  return refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

double ref_pr(void* gp, int max_iters, double epsilon, float* scores_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  auto t0 = std::chrono::steady_clock::now();
  pvector<ScoreT> scores = PageRankPullGS(g, max_iters, epsilon);
  auto t1 = std::chrono::steady_clock::now();
  if (scores_out) std::copy(scores.begin(), scores.end(), scores_out);
  return refshim::Seconds(t0, t1);
}

}  // extern "C"
