// Test infrastructure only: the reference's Hybrid / OrderedCount
// (src/tc.cc:52-100) behind a C-ABI for parity tests and the CPU baseline.
#define main gapbs_reference_main
#include "tc.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                    const int64_t* ioff, const int32_t* inbr) {
  return refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

// mode 0: Hybrid (relabel if worth it); 1: OrderedCount without relabel.
double ref_tc(void* gp, int mode, uint64_t* count_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  auto t0 = std::chrono::steady_clock::now();
  size_t c = mode == 0 ? Hybrid(g) : OrderedCount(g);
  auto t1 = std::chrono::steady_clock::now();
  *count_out = c;
  return refshim::Seconds(t0, t1);
}

int ref_tc_worth_relabelling(void* gp) { return WorthRelabelling(*static_cast<Graph*>(gp)) ? 1 : 0; }

}  // extern "C"
