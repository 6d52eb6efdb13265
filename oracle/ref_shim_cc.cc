// Test infrastructure only: the reference's Afforest (src/cc.cc:95-151)
// behind a C-ABI for parity tests and the CPU baseline.
#define main gapbs_reference_main
#include "cc.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                    const int64_t* ioff, const int32_t* inbr) {
  return refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

double ref_cc(void* gp, int32_t* comp_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  auto t0 = std::chrono::steady_clock::now();
  pvector<NodeID> comp = Afforest(g);
  auto t1 = std::chrono::steady_clock::now();
  if (comp_out) std::copy(comp.begin(), comp.end(), comp_out);
  return refshim::Seconds(t0, t1);
}

}  // extern "C"
