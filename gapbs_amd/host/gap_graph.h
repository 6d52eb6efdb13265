// Host CSR graph container.
//
// Reference: CSRGraph / NodeWeight (src/graph.h:29-55, 98-273).  The
// reference keeps an index of N+1 *pointers* into the neighbour array; this
// container keeps int64 offsets instead, which is the layout that is copied
// verbatim into HBM and into the .sg/.wsg files (src/writer.h:39-69).
// Semantics that callers depend on are kept:
//   * an undirected graph shares one array for in- and out-neighbours and
//     reports num_edges = entries/2 (graph.h:131-143);
//   * num_edges_directed() = 2*num_edges for undirected graphs (201-203);
//   * out_neigh(v, start) clamps start to the degree (graph.h:104-117),
//     which Afforest relies on (src/cc.cc:108, 129);
//   * the stats line format (graph.h:223-230).
#pragma once

#include <iostream>
#include <sstream>
#include <string>

#include "gap_core.h"

namespace gap {

// Weighted neighbour: 8 bytes, {v, w}, the .wsg on-disk element.
struct WNode {
  NodeID v;
  WeightT w;
};

inline NodeID DestId(NodeID d) { return d; }
inline NodeID DestId(const WNode& d) { return d.v; }

template <class DestT>
class CSRGraphT {
 public:
  using Dest = DestT;

  struct Neigh {
    const DestT* b;
    const DestT* e;
    const DestT* begin() const { return b; }
    const DestT* end() const { return e; }
    int64_t size() const { return e - b; }
  };

  CSRGraphT() = default;

  // Undirected graph: out == in.
  CSRGraphT(int64_t n, Buf<Offset>&& off, Buf<DestT>&& dst)
      : directed_(false), n_(n), out_off_(std::move(off)), out_dst_(std::move(dst)) {
    num_edges_ = (out_off_[n_] - out_off_[0]) / 2;
  }

  // Directed graph with its inverse.
  CSRGraphT(int64_t n, Buf<Offset>&& off, Buf<DestT>&& dst, Buf<Offset>&& in_off,
            Buf<DestT>&& in_dst)
      : directed_(true),
        n_(n),
        out_off_(std::move(off)),
        out_dst_(std::move(dst)),
        in_off_(std::move(in_off)),
        in_dst_(std::move(in_dst)) {
    num_edges_ = out_off_[n_] - out_off_[0];
  }

  CSRGraphT(CSRGraphT&&) = default;
  CSRGraphT& operator=(CSRGraphT&&) = default;

  bool directed() const { return directed_; }
  int64_t num_nodes() const { return n_; }
  int64_t num_edges() const { return num_edges_; }
  int64_t num_edges_directed() const { return directed_ ? num_edges_ : 2 * num_edges_; }
  // Number of stored out-entries (== num_edges_directed() for graphs built
  // here; kept separate because the reference derives the latter).
  int64_t out_entries() const { return out_off_[n_]; }

  int64_t out_degree(NodeID v) const { return out_off_[v + 1] - out_off_[v]; }
  int64_t in_degree(NodeID v) const {
    const Offset* o = in_offsets();
    return o[v + 1] - o[v];
  }

  Neigh out_neigh(NodeID v, uint64_t start = 0) const {
    return Clamp(out_off_.data(), out_dst_.data(), v, start);
  }
  Neigh in_neigh(NodeID v, uint64_t start = 0) const {
    return Clamp(in_offsets(), in_dests(), v, start);
  }

  const Offset* out_offsets() const { return out_off_.data(); }
  const DestT* out_dests() const { return out_dst_.data(); }
  const Offset* in_offsets() const { return directed_ ? in_off_.data() : out_off_.data(); }
  const DestT* in_dests() const { return directed_ ? in_dst_.data() : out_dst_.data(); }
  int64_t in_entries() const { return in_offsets()[n_]; }

  std::string StatsLine() const {
    std::ostringstream os;
    os << "Graph has " << n_ << " nodes and " << num_edges_ << " " << (directed_ ? "" : "un")
       << "directed edges for degree: " << (n_ ? num_edges_ / n_ : 0);
    return os.str();
  }
  void PrintStats() const { std::cout << StatsLine() << std::endl; }

 private:
  static Neigh Clamp(const Offset* off, const DestT* dst, NodeID v, uint64_t start) {
    const DestT* b = dst + off[v];
    const DestT* e = dst + off[v + 1];
    uint64_t deg = static_cast<uint64_t>(e - b);
    if (start > deg) start = deg;
    return Neigh{b + start, e};
  }

  bool directed_ = false;
  int64_t n_ = 0;
  int64_t num_edges_ = 0;
  Buf<Offset> out_off_;
  Buf<DestT> out_dst_;
  Buf<Offset> in_off_;
  Buf<DestT> in_dst_;
};

using Graph = CSRGraphT<NodeID>;
using WGraph = CSRGraphT<WNode>;

}  // namespace gap
