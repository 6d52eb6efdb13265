// Test infrastructure only: exposes the reference's graph builder
// (BuilderBase::MakeGraph, src/builder.h:343-364) and SourcePicker
// (src/benchmark.h:46-68) so tests can compare libgaphost's generator and
// root sequence with the reference's on the same flags.
#include "benchmark.h"
#include "builder.h"
#include "command_line.h"

#include <getopt.h>

#include <string>
#include <vector>

#include "ref_shim_common.h"

namespace {
template <class GraphT>
void CopyOut(const GraphT& g, int64_t* off, void* dst, int64_t* ioff, void* idst) {
  typedef typename std::remove_pointer<decltype(g.out_neigh(0).begin())>::type DestT;
  pvector<SGOffset> o = g.VertexOffsets(false);
  for (int64_t i = 0; i <= g.num_nodes(); i++) off[i] = o[i];
  std::memcpy(dst, g.out_neigh(0).begin(), sizeof(DestT) * o[g.num_nodes()]);
  if (g.directed() && ioff) {
    pvector<SGOffset> io = g.VertexOffsets(true);
    for (int64_t i = 0; i <= g.num_nodes(); i++) ioff[i] = io[i];
    std::memcpy(idst, g.in_neigh(0).begin(), sizeof(DestT) * io[g.num_nodes()]);
  }
}
}  // namespace

extern "C" {

// argv[0] is ignored; the rest are the reference's own flags (-g, -u, -k, -f, -s).
void* ref_build(int argc, const char** argv, int weighted) {
  std::vector<std::string> store(argv, argv + argc);
  std::vector<char*> args;
  for (auto& s : store) args.push_back(&s[0]);
  args.push_back(nullptr);
  optind = 0;
  CLBase cli(argc, args.data(), "ref_build");
  if (!cli.ParseArgs()) return nullptr;
  if (weighted) {
    WeightedBuilder b(cli);
    return new WGraph(b.MakeGraph());
  }
  Builder b(cli);
  return new Graph(b.MakeGraph());
}

void ref_build_free(void* g, int weighted) {
  if (weighted)
    delete static_cast<WGraph*>(g);
  else
    delete static_cast<Graph*>(g);
}

void ref_build_shape(void* gp, int weighted, int64_t* nodes, int64_t* entries, int* directed) {
  if (weighted) {
    WGraph& g = *static_cast<WGraph*>(gp);
    *nodes = g.num_nodes(); *entries = g.num_edges_directed(); *directed = g.directed();
  } else {
    Graph& g = *static_cast<Graph*>(gp);
    *nodes = g.num_nodes(); *entries = g.num_edges_directed(); *directed = g.directed();
  }
}

void ref_build_copy(void* gp, int weighted, int64_t* off, void* dst, int64_t* ioff, void* idst) {
  if (weighted)
    CopyOut(*static_cast<WGraph*>(gp), off, dst, ioff, idst);
  else
    CopyOut(*static_cast<Graph*>(gp), off, dst, ioff, idst);
}

// First k roots of SourcePicker(g, given) on a graph given as arrays.
void ref_pick_sources(int64_t n, const int64_t* off, const int32_t* nbr, int directed,
                      const int64_t* ioff, const int32_t* inbr, int64_t given, int k, int32_t* out) {
  Graph* g = refshim::MakeGraph<NodeID>(n, off, nbr, directed, ioff, inbr);
  SourcePicker<Graph> sp(*g, given);
  for (int i = 0; i < k; i++) out[i] = sp.PickNext();
  delete g;
}

}  // extern "C"
