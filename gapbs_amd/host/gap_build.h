// CSR construction from an edge list, and relabelling by degree.
//
// Reference: BuilderBase (src/builder.h).  The result must be the CSR the
// reference produces for the same edge list, because every kernel result is
// compared position by position:
//   * vertex count = largest endpoint + 1 (FindMaxNodeID, builder.h:69-78,
//     used whenever the count is not known, i.e. always, builder.h:321-322);
//   * symmetrised graphs store both directions of every edge, with the same
//     weight (MakeCSR, builder.h:300-315);
//   * every neighbour list is sorted, duplicates are dropped keeping the
//     smallest weight (sort by (v, w), unique by v) and self-loops removed
//     (SquishCSR, builder.h:141-170);
//   * directed graphs also get the squished transpose (builder.h:172-186);
//   * RelabelByDegree orders vertices by descending (degree, old id) and
//     re-sorts every list (builder.h:367-402).
#pragma once

#include <algorithm>
#include <vector>
#include <functional>
#include <utility>

#include "gap_core.h"
#include "gap_gen.h"
#include "gap_graph.h"

namespace gap {

inline int64_t FindNumNodes(const Buf<Edge>& el) {
  NodeID mx = -1;
  const int64_t m = static_cast<int64_t>(el.size());
#pragma omp parallel for reduction(max : mx)
  for (int64_t e = 0; e < m; e++) mx = std::max(mx, std::max(el[e].u, el[e].v));
  return static_cast<int64_t>(mx) + 1;
}

namespace detail {

template <class DestT>
DestT Dest(NodeID v, const WeightT* w, int64_t e);
template <>
inline NodeID Dest<NodeID>(NodeID v, const WeightT*, int64_t) {
  return v;
}
template <>
inline WNode Dest<WNode>(NodeID v, const WeightT* w, int64_t e) {
  return WNode{v, w ? w[e] : 1};
}

inline bool DestLess(NodeID a, NodeID b) { return a < b; }
inline bool DestLess(const WNode& a, const WNode& b) { return a.v == b.v ? a.w < b.w : a.v < b.v; }

// Scatters the edge list into an unsquished CSR.  forward: u -> v entries;
// backward: v -> u entries (both for a symmetrised graph).
template <class DestT>
void Scatter(const Buf<Edge>& el, const WeightT* w, int64_t n, bool forward, bool backward,
             Buf<Offset>* off, Buf<DestT>* dst) {
  const int64_t m = static_cast<int64_t>(el.size());
  Buf<int64_t> deg(static_cast<size_t>(n), 0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    if (forward) {
#pragma omp atomic
      deg[el[e].u]++;
    }
    if (backward) {
#pragma omp atomic
      deg[el[e].v]++;
    }
  }
  *off = PrefixSum(deg.data(), n);
  dst->reset(static_cast<size_t>((*off)[n]));
  Buf<int64_t>& cursor = deg;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) cursor[i] = (*off)[i];
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    const NodeID u = el[e].u, v = el[e].v;
    if (forward) {
      int64_t p;
#pragma omp atomic capture
      p = cursor[u]++;
      (*dst)[p] = Dest<DestT>(v, w, e);
    }
    if (backward) {
      int64_t p;
#pragma omp atomic capture
      p = cursor[v]++;
      (*dst)[p] = Dest<DestT>(u, w, e);
    }
  }
}

// Sorts each list, drops duplicates (by destination id) and self-loops.
template <class DestT>
void Squish(int64_t n, Buf<Offset>* off, Buf<DestT>* dst) {
  Buf<int64_t> keep(static_cast<size_t>(n));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < n; u++) {
    DestT* b = dst->data() + (*off)[u];
    DestT* e = dst->data() + (*off)[u + 1];
    std::sort(b, e, [](const DestT& x, const DestT& y) { return DestLess(x, y); });
    DestT* out = b;
    for (DestT* it = b; it != e; ++it) {
      const NodeID v = DestId(*it);
      if (v == static_cast<NodeID>(u)) continue;
      if (out != b && DestId(*(out - 1)) == v) continue;
      *out++ = *it;
    }
    keep[u] = out - b;
  }
  Buf<Offset> noff = PrefixSum(keep.data(), n);
  Buf<DestT> ndst(static_cast<size_t>(noff[n]));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < n; u++)
    std::copy(dst->data() + (*off)[u], dst->data() + (*off)[u] + keep[u], ndst.data() + noff[u]);
  *off = std::move(noff);
  *dst = std::move(ndst);
}

}  // namespace detail

// Builds the squished CSR.  `w` (nullable) holds one weight per edge-list
// entry and is only read for weighted graphs.  The edge list is consumed.
template <class DestT>
CSRGraphT<DestT> BuildFromEdgeList(Buf<Edge>&& el_in, const WeightT* w, bool symmetrize) {
  Buf<Edge> el = std::move(el_in);
  const int64_t n = FindNumNodes(el);
  Buf<Offset> off, in_off;
  Buf<DestT> dst, in_dst;
  if (symmetrize) {
    detail::Scatter<DestT>(el, w, n, true, true, &off, &dst);
  } else {
    detail::Scatter<DestT>(el, w, n, true, false, &off, &dst);
    detail::Scatter<DestT>(el, w, n, false, true, &in_off, &in_dst);
  }
  el.reset(0);
  detail::Squish(n, &off, &dst);
  if (symmetrize) return CSRGraphT<DestT>(n, std::move(off), std::move(dst));
  detail::Squish(n, &in_off, &in_dst);
  return CSRGraphT<DestT>(n, std::move(off), std::move(dst), std::move(in_off), std::move(in_dst));
}

// New id of every vertex when ordered by descending (degree, old id).
inline Buf<NodeID> DegreeOrder(const Graph& g) {
  const int64_t n = g.num_nodes();
  std::vector<std::pair<int64_t, NodeID>> order(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; v++)
    order[v] = std::make_pair(g.out_degree(static_cast<NodeID>(v)), static_cast<NodeID>(v));
  std::sort(order.begin(), order.end(), std::greater<std::pair<int64_t, NodeID>>());
  Buf<NodeID> new_id(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) new_id[order[i].second] = static_cast<NodeID>(i);
  return new_id;
}

// Returns the graph relabelled by descending degree (undirected only).
inline Graph RelabelByDegree(const Graph& g) {
  if (g.directed()) throw Error(-11, "Cannot relabel directed graph");
  const int64_t n = g.num_nodes();
  Buf<NodeID> new_id = DegreeOrder(g);
  Buf<int64_t> deg(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; v++) deg[new_id[v]] = g.out_degree(static_cast<NodeID>(v));
  Buf<Offset> off = PrefixSum(deg.data(), n);
  Buf<NodeID> dst(static_cast<size_t>(off[n]));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < n; u++) {
    const NodeID nu = new_id[u];
    NodeID* out = dst.data() + off[nu];
    int64_t k = 0;
    for (NodeID v : g.out_neigh(static_cast<NodeID>(u))) out[k++] = new_id[v];
    std::sort(out, out + k);
  }
  return Graph(n, std::move(off), std::move(dst));
}

// Drops the weights of a weighted graph (same topology).
inline Graph StripWeights(const WGraph& g) {
  const int64_t n = g.num_nodes();
  auto copy_side = [n](const Offset* o, const WNode* d, Buf<Offset>* off, Buf<NodeID>* dst) {
    off->reset(static_cast<size_t>(n) + 1);
    std::copy(o, o + n + 1, off->data());
    dst->reset(static_cast<size_t>(o[n]));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < o[n]; i++) (*dst)[i] = d[i].v;
  };
  Buf<Offset> off, ioff;
  Buf<NodeID> dst, idst;
  copy_side(g.out_offsets(), g.out_dests(), &off, &dst);
  if (!g.directed()) return Graph(n, std::move(off), std::move(dst));
  copy_side(g.in_offsets(), g.in_dests(), &ioff, &idst);
  return Graph(n, std::move(off), std::move(dst), std::move(ioff), std::move(idst));
}

// Disjoint union of unweighted graphs of one kind (all undirected or all
// directed): part i's vertices are shifted by the vertex counts of parts
// 0 .. i-1, its lists stay sorted and duplicate-free.  No reference
// counterpart; used to run many small graphs in one batched traversal.
inline Graph PackGraphs(const std::vector<const Graph*>& parts, std::vector<int64_t>* bounds) {
  if (parts.empty()) throw Error(-2, "no graphs to pack");
  const bool directed = parts[0]->directed();
  bounds->assign(1, 0);
  int64_t n = 0, e_out = 0, e_in = 0;
  for (const Graph* p : parts) {
    if (p->directed() != directed) throw Error(-2, "cannot pack directed with undirected graphs");
    n += p->num_nodes();
    e_out += p->out_offsets()[p->num_nodes()];
    e_in += p->in_offsets()[p->num_nodes()];
    bounds->push_back(n);
  }
  if (n >= (int64_t(1) << 31)) throw Error(-2, "packed graph exceeds int32 vertex ids");
  auto pack_side = [&](bool in, int64_t entries, Buf<Offset>* off, Buf<NodeID>* dst) {
    off->reset(static_cast<size_t>(n) + 1);
    dst->reset(static_cast<size_t>(entries));
    int64_t vbase = 0, ebase = 0;
    for (const Graph* p : parts) {
      const Offset* o = in ? p->in_offsets() : p->out_offsets();
      const NodeID* d = in ? p->in_dests() : p->out_dests();
      const int64_t pn = p->num_nodes(), pe = o[pn];
      for (int64_t v = 0; v < pn; v++) (*off)[vbase + v] = ebase + o[v];
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < pe; k++) (*dst)[ebase + k] = static_cast<NodeID>(d[k] + vbase);
      vbase += pn;
      ebase += pe;
    }
    (*off)[n] = ebase;
  };
  Buf<Offset> off, ioff;
  Buf<NodeID> dst, idst;
  pack_side(false, e_out, &off, &dst);
  if (!directed) return Graph(n, std::move(off), std::move(dst));
  pack_side(true, e_in, &ioff, &idst);
  return Graph(n, std::move(off), std::move(dst), std::move(ioff), std::move(idst));
}

}  // namespace gap
