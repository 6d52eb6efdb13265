// Test infrastructure only: the reference's DOBFS (src/bfs.cc:122-179),
// compiled from the reference source, behind a C-ABI for parity tests and
// for the CPU baseline.
#define main gapbs_reference_main
#include "bfs.cc"
#undef main

#include "ref_shim_common.h"

extern "C" {

void* ref_graph_new(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                    const int64_t* ioff, const int32_t* inbr) {
  return refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

// Runs DOBFS from each root; returns the summed kernel seconds.  parent_out
// (nullable) receives k * n parents, root-major.
double ref_bfs(void* gp, const int32_t* roots, int k, int32_t* parent_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  double total = 0;
  for (int i = 0; i < k; i++) {
    auto t0 = std::chrono::steady_clock::now();
    pvector<NodeID> parent = DOBFS(g, roots[i]);
    auto t1 = std::chrono::steady_clock::now();
    total += refshim::Seconds(t0, t1);
    if (parent_out) std::copy(parent.begin(), parent.end(), parent_out + int64_t(i) * g.num_nodes());
  }
  return total;
}

int ref_bfs_verify(void* gp, int32_t root, const int32_t* parent) {
  const Graph& g = *static_cast<Graph*>(gp);
  pvector<NodeID> p(g.num_nodes());
  std::copy(parent, parent + g.num_nodes(), p.begin());
  return BFSVerifier(g, root, p) ? 1 : 0;
}

}  // extern "C"
