// Test infrastructure only: the reference's Brandes (src/bc.cc:95-150)
// behind a C-ABI for parity tests and the CPU baseline.  Roots come from
// the reference's SourcePicker (given == -1) or are all `given`.
#define main gapbs_reference_main
#include "bc.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                    const int64_t* ioff, const int32_t* inbr) {
  return refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

double ref_bc(void* gp, int num_iters, int64_t given, float* scores_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  SourcePicker<Graph> sp(g, given);
  auto t0 = std::chrono::steady_clock::now();
  pvector<ScoreT> scores = Brandes(g, sp, num_iters);
  auto t1 = std::chrono::steady_clock::now();
  if (scores_out) std::copy(scores.begin(), scores.end(), scores_out);
  return refshim::Seconds(t0, t1);
}

}  // extern "C"
