// Test infrastructure only: the reference's DeltaStep (src/sssp.cc:87-153)
// behind a C-ABI for parity tests and the CPU baseline.
#define main gapbs_reference_main
#include "sssp.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const void* wnbr, int directed,
                    const int64_t* ioff, const void* iwnbr) {
  return refshim::MakeGraph<WNode>(n, off, wnbr, directed, ioff, iwnbr);
}

void ref_graph_free(void* g) { delete static_cast<WGraph*>(g); }

double ref_sssp(void* gp, const int32_t* roots, int k, int32_t delta, int32_t* dist_out) {
  const WGraph& g = *static_cast<WGraph*>(gp);
  double total = 0;
  for (int i = 0; i < k; i++) {
    auto t0 = std::chrono::steady_clock::now();
    pvector<WeightT> dist = DeltaStep(g, roots[i], delta);
    auto t1 = std::chrono::steady_clock::now();
    total += refshim::Seconds(t0, t1);
    if (dist_out) std::copy(dist.begin(), dist.end(), dist_out + int64_t(i) * g.num_nodes());
  }
  return total;
}

}  // extern "C"
