// Per-kernel verifiers and analysis printers (the `-v` and `-a` flags).
//
// Each verifier is a serial CPU check of one kernel result, restated from
// the reference:
//   BFSVerifier  src/bfs.cc:201-247   serial BFS depths; parent edges checked
//   SSSPVerifier src/sssp.cc:164-194  serial binary-heap Dijkstra, exact
//   PRVerifier   src/pr.cc:78-94      one push iteration, L1 error < tol
//   CCVerifier   src/cc.cc:177-217    BFS per label, no label crossing
//   BCVerifier   src/bc.cc:169-230    serial Brandes, |diff| <= FLT_EPSILON
//   TCVerifier   src/tc.cc:109-128    set_intersection count / 6
// Printers: bfs.cc:182-193, sssp.cc:156-160, pr.cc:64-73, cc.cc:154-170,
// bc.cc:153-161, tc.cc:103-105.
#pragma once

#include <cmath>
#include <iostream>
#include <limits>
#include <queue>
#include <unordered_map>
#include <vector>

#include "gap_core.h"
#include "gap_graph.h"
#include "gap_harness.h"

namespace gap {

constexpr WeightT kDistInf = std::numeric_limits<WeightT>::max() / 2;

// ---- BFS -------------------------------------------------------------------

inline std::vector<int32_t> SerialBFSDepths(const Graph& g, NodeID source) {
  std::vector<int32_t> depth(static_cast<size_t>(g.num_nodes()), -1);
  std::vector<NodeID> order;
  order.reserve(static_cast<size_t>(g.num_nodes()));
  depth[source] = 0;
  order.push_back(source);
  for (size_t i = 0; i < order.size(); i++) {
    const NodeID u = order[i];
    for (NodeID v : g.out_neigh(u))
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        order.push_back(v);
      }
  }
  return depth;
}

inline bool BFSVerifier(const Graph& g, NodeID source, const NodeID* parent) {
  const std::vector<int32_t> depth = SerialBFSDepths(g, source);
  for (NodeID u = 0; u < g.num_nodes(); u++) {
    if (depth[u] != -1 && parent[u] != -1) {
      if (u == source) {
        if (parent[u] != u || depth[u] != 0) {
          std::cout << "Source wrong" << std::endl;
          return false;
        }
        continue;
      }
      bool found = false;
      for (NodeID v : g.in_neigh(u)) {
        if (v != parent[u]) continue;
        if (depth[v] != depth[u] - 1) {
          std::cout << "Wrong depths for " << u << " & " << v << std::endl;
          return false;
        }
        found = true;
        break;
      }
      if (!found) {
        std::cout << "Couldn't find edge from " << parent[u] << " to " << u << std::endl;
        return false;
      }
    } else if (depth[u] != parent[u]) {
      std::cout << "Reachability mismatch" << std::endl;
      return false;
    }
  }
  return true;
}

inline void PrintBFSStats(const Graph& g, const NodeID* parent) {
  int64_t tree = 0, edges = 0;
  for (NodeID n = 0; n < g.num_nodes(); n++)
    if (parent[n] >= 0) {
      edges += g.out_degree(n);
      tree++;
    }
  std::cout << "BFS Tree has " << tree << " nodes and " << edges << " edges" << std::endl;
}

// ---- SSSP ------------------------------------------------------------------

inline std::vector<WeightT> SerialDijkstra(const WGraph& g, NodeID source) {
  std::vector<WeightT> dist(static_cast<size_t>(g.num_nodes()), kDistInf);
  using Item = std::pair<WeightT, NodeID>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  dist[source] = 0;
  pq.emplace(0, source);
  while (!pq.empty()) {
    const WeightT d = pq.top().first;
    const NodeID u = pq.top().second;
    pq.pop();
    if (d != dist[u]) continue;
    for (const WNode& wn : g.out_neigh(u))
      if (d + wn.w < dist[wn.v]) {
        dist[wn.v] = d + wn.w;
        pq.emplace(d + wn.w, wn.v);
      }
  }
  return dist;
}

inline bool SSSPVerifier(const WGraph& g, NodeID source, const WeightT* dist) {
  const std::vector<WeightT> oracle = SerialDijkstra(g, source);
  bool ok = true;
  for (NodeID n = 0; n < g.num_nodes(); n++)
    if (dist[n] != oracle[n]) {
      std::cout << n << ": " << dist[n] << " != " << oracle[n] << std::endl;
      ok = false;
    }
  return ok;
}

inline void PrintSSSPStats(const WGraph& g, const WeightT* dist) {
  int64_t reached = 0;
  for (NodeID n = 0; n < g.num_nodes(); n++) reached += dist[n] != kDistInf;
  std::cout << "SSSP Tree reaches " << reached << " nodes" << std::endl;
}

// ---- PageRank --------------------------------------------------------------

constexpr float kDamp = 0.85f;

inline bool PRVerifier(const Graph& g, const float* scores, double target_error) {
  const float base = (1.0f - kDamp) / g.num_nodes();
  std::vector<float> incoming(static_cast<size_t>(g.num_nodes()), 0.0f);
  for (NodeID u = 0; u < g.num_nodes(); u++) {
    const float contrib = scores[u] / g.out_degree(u);
    for (NodeID v : g.out_neigh(u)) incoming[v] += contrib;
  }
  double error = 0;
  for (NodeID n = 0; n < g.num_nodes(); n++) error += std::fabs(base + kDamp * incoming[n] - scores[n]);
  PrintTime("Total Error", error);
  return error < target_error;
}

inline void PrintTopScores(const Graph& g, const float* scores) {
  std::vector<std::pair<NodeID, float>> kv(static_cast<size_t>(g.num_nodes()));
  for (NodeID n = 0; n < g.num_nodes(); n++) kv[n] = std::make_pair(n, scores[n]);
  for (const auto& p : TopK(kv, 5)) std::cout << p.second << ":" << p.first << std::endl;
}

// ---- CC --------------------------------------------------------------------

inline bool CCVerifier(const Graph& g, const NodeID* comp) {
  std::unordered_map<NodeID, NodeID> label_to_source;
  for (NodeID n = 0; n < g.num_nodes(); n++) label_to_source[comp[n]] = n;
  std::vector<uint8_t> visited(static_cast<size_t>(g.num_nodes()), 0);
  std::vector<NodeID> frontier;
  frontier.reserve(static_cast<size_t>(g.num_nodes()));
  for (const auto& ls : label_to_source) {
    const NodeID label = ls.first, source = ls.second;
    frontier.clear();
    frontier.push_back(source);
    visited[source] = 1;
    for (size_t i = 0; i < frontier.size(); i++) {
      const NodeID u = frontier[i];
      auto visit = [&](NodeID v) {
        if (comp[v] != label) return false;
        if (!visited[v]) {
          visited[v] = 1;
          frontier.push_back(v);
        }
        return true;
      };
      for (NodeID v : g.out_neigh(u))
        if (!visit(v)) return false;
      if (g.directed())
        for (NodeID v : g.in_neigh(u))
          if (!visit(v)) return false;
    }
  }
  for (NodeID n = 0; n < g.num_nodes(); n++)
    if (!visited[n]) return false;
  return true;
}

inline void PrintCompStats(const Graph& g, const NodeID* comp) {
  std::cout << std::endl;
  std::unordered_map<NodeID, NodeID> count;
  for (NodeID n = 0; n < g.num_nodes(); n++) count[comp[n]] += 1;
  std::vector<std::pair<NodeID, NodeID>> kv(count.begin(), count.end());
  auto top = TopK(kv, 5);
  std::cout << top.size() << " biggest clusters" << std::endl;
  for (const auto& p : top) std::cout << p.second << ":" << p.first << std::endl;
  std::cout << "There are " << count.size() << " components" << std::endl;
}

// ---- BC --------------------------------------------------------------------

// Serial Brandes over `roots`, max-normalised; float scores, double counts.
inline std::vector<float> SerialBrandes(const Graph& g, const std::vector<NodeID>& roots) {
  const int64_t n = g.num_nodes();
  std::vector<float> scores(static_cast<size_t>(n), 0.0f);
  for (NodeID source : roots) {
    std::vector<int32_t> depth(static_cast<size_t>(n), -1);
    std::vector<double> sigma(static_cast<size_t>(n), 0.0);
    std::vector<NodeID> order;
    order.reserve(static_cast<size_t>(n));
    depth[source] = 0;
    sigma[source] = 1;
    order.push_back(source);
    for (size_t i = 0; i < order.size(); i++) {
      const NodeID u = order[i];
      for (NodeID v : g.out_neigh(u)) {
        if (depth[v] == -1) {
          depth[v] = depth[u] + 1;
          order.push_back(v);
        }
        if (depth[v] == depth[u] + 1) sigma[v] += sigma[u];
      }
    }
    std::vector<std::vector<NodeID>> levels;
    for (NodeID v = 0; v < n; v++)
      if (depth[v] != -1) {
        if (depth[v] >= static_cast<int32_t>(levels.size())) levels.resize(depth[v] + 1);
        levels[depth[v]].push_back(v);
      }
    std::vector<float> delta(static_cast<size_t>(n), 0.0f);
    for (int d = static_cast<int>(levels.size()) - 1; d >= 0; d--)
      for (NodeID u : levels[d]) {
        for (NodeID v : g.out_neigh(u))
          if (depth[v] == depth[u] + 1) delta[u] += (sigma[u] / sigma[v]) * (1 + delta[v]);
        scores[u] += delta[u];
      }
  }
  float biggest = 0;
  for (float s : scores) biggest = std::max(biggest, s);
  for (float& s : scores) s = s / biggest;
  return scores;
}

inline bool BCVerifier(const Graph& g, SourcePicker<Graph>& sp, int num_iters, const float* scores) {
  std::vector<NodeID> roots;
  for (int i = 0; i < num_iters; i++) roots.push_back(sp.PickNext());
  const std::vector<float> oracle = SerialBrandes(g, roots);
  bool ok = true;
  for (NodeID n = 0; n < g.num_nodes(); n++) {
    const float d = std::fabs(scores[n] - oracle[n]);
    if (d > std::numeric_limits<float>::epsilon()) {
      std::cout << n << ": " << oracle[n] << " != " << scores[n] << "(" << d << ")" << std::endl;
      ok = false;
    }
  }
  return ok;
}

// ---- TC --------------------------------------------------------------------

inline bool TCVerifier(const Graph& g, uint64_t test_total) {
  uint64_t total = 0;
  std::vector<NodeID> inter;
  inter.reserve(static_cast<size_t>(g.num_nodes()));
  for (NodeID u = 0; u < g.num_nodes(); u++)
    for (NodeID v : g.out_neigh(u)) {
      inter.clear();
      auto nu = g.out_neigh(u), nv = g.out_neigh(v);
      std::set_intersection(nu.begin(), nu.end(), nv.begin(), nv.end(), std::back_inserter(inter));
      total += inter.size();
    }
  total /= 6;
  if (total != test_total) std::cout << total << " != " << test_total << std::endl;
  return total == test_total;
}

inline void PrintTriangleStats(const Graph&, uint64_t total) {
  std::cout << total << " triangles" << std::endl;
}

}  // namespace gap
