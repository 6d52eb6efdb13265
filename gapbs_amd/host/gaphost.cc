// C-ABI over the host graph layer (declared in include/gaphost.h).
#include "gaphost.h"

#include <cstring>
#include <string>

#include "gap_build.h"
#include "gap_gen.h"
#include "gap_graph.h"
#include "gap_harness.h"
#include "gap_sgio.h"
#include "gap_verify.h"

using namespace gap;

struct gh_graph {
  bool weighted = false;
  Graph g;
  WGraph wg;
};

namespace {

thread_local std::string g_error;

template <class F>
int Guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    g_error = e.what();
    return e.code < 0 ? e.code : -1;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

gh_graph* Wrap(Graph&& g) {
  auto* h = new gh_graph;
  h->g = std::move(g);
  return h;
}
gh_graph* Wrap(WGraph&& g) {
  auto* h = new gh_graph;
  h->weighted = true;
  h->wg = std::move(g);
  return h;
}

const Graph& Unweighted(const gh_graph* h) {
  if (h->weighted) throw Error(-5, "operation needs an unweighted graph");
  return h->g;
}
const WGraph& Weighted(const gh_graph* h) {
  if (!h->weighted) throw Error(-5, "operation needs a weighted graph");
  return h->wg;
}

}  // namespace

extern "C" {

int gh_generate(int kind, int scale, int degree, int64_t seed, int weighted, gh_graph** out) {
  return Guard([&] {
    Buf<WeightT> w;
    Buf<Edge> el;
    if (kind == GH_GRID) {
      GridSpec sp;
      sp.side = scale;
      sp.seed = seed;
      el = MakeGridEL(sp, &w);
    } else if (kind == GH_UNIFORM) {
      el = MakeUniformEL(scale, degree, seed);
    } else if (kind == GH_KRONECKER) {
      el = MakeKroneckerEL(scale, degree, seed);
    } else {
      throw Error(-3, "unknown generator kind");
    }
    if (weighted) {
      if (kind != GH_GRID) w = MakeWeights(static_cast<int64_t>(el.size()), seed);
      *out = Wrap(BuildFromEdgeList<WNode>(std::move(el), w.data(), true));
    } else {
      *out = Wrap(BuildFromEdgeList<NodeID>(std::move(el), nullptr, true));
    }
    return 0;
  });
}

int gh_permutation(int64_t n, int64_t seed, int32_t* perm) {
  return Guard([&] {
    if (n < 0 || n > (int64_t(1) << 31) || (n > 0 && !perm)) throw Error(-2, "bad permutation arguments");
    for (int64_t i = 0; i < n; i++) perm[i] = static_cast<int32_t>(i);
    std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
    std::shuffle(perm, perm + n, rng);
    return 0;
  });
}

int gh_generate_edges(int kind, int scale, int degree, int64_t seed, int weighted, int64_t* m, int32_t* u,
                      int32_t* v, int32_t* w) {
  return Guard([&] {
    if (!m) throw Error(-2, "null edge count");
    Buf<WeightT> wt;
    Buf<Edge> el;
    if (kind == GH_GRID) {
      GridSpec sp;
      sp.side = scale;
      sp.seed = seed;
      el = MakeGridEL(sp, &wt);
    } else if (kind == GH_UNIFORM) {
      el = MakeUniformEL(scale, degree, seed);
    } else if (kind == GH_KRONECKER) {
      el = MakeKroneckerEL(scale, degree, seed);
    } else {
      throw Error(-3, "unknown generator kind");
    }
    const int64_t count = static_cast<int64_t>(el.size());
    if (u) {
      if (*m < count || !v || (weighted && !w)) throw Error(-2, "edge arrays too small or missing");
      if (weighted && kind != GH_GRID) wt = MakeWeights(count, seed);
#pragma omp parallel for schedule(static)
      for (int64_t e = 0; e < count; e++) {
        u[e] = el[e].u;
        v[e] = el[e].v;
        if (weighted) w[e] = wt[e];
      }
    }
    *m = count;
    return 0;
  });
}

int gh_build_edges(const int32_t* u, const int32_t* v, const int32_t* w, int64_t m, int symmetrize,
                   int weighted, gh_graph** out) {
  return Guard([&] {
    Buf<Edge> el(static_cast<size_t>(m));
    for (int64_t i = 0; i < m; i++) {
      if (u[i] < 0 || v[i] < 0) throw Error(-2, "negative vertex id in edge list");
      el[i] = Edge{u[i], v[i]};
    }
    if (weighted)
      *out = Wrap(BuildFromEdgeList<WNode>(std::move(el), w, symmetrize != 0));
    else
      *out = Wrap(BuildFromEdgeList<NodeID>(std::move(el), nullptr, symmetrize != 0));
    return 0;
  });
}

int gh_read(const char* path, int weighted, gh_graph** out) {
  return Guard([&] {
    if (weighted)
      *out = Wrap(ReadSerialized<WNode>(path));
    else
      *out = Wrap(ReadSerialized<NodeID>(path));
    return 0;
  });
}

int gh_write(const gh_graph* g, const char* path) {
  return Guard([&] {
    if (g->weighted)
      WriteSerialized(g->wg, path);
    else
      WriteSerialized(g->g, path);
    return 0;
  });
}

void gh_free(gh_graph* g) { delete g; }

int gh_shape(const gh_graph* g, int64_t* nodes, int64_t* entries, int* directed, int* weighted) {
  return Guard([&] {
    const bool w = g->weighted;
    *nodes = w ? g->wg.num_nodes() : g->g.num_nodes();
    *entries = w ? g->wg.num_edges_directed() : g->g.num_edges_directed();
    *directed = w ? g->wg.directed() : g->g.directed();
    *weighted = w;
    return 0;
  });
}

int gh_stats_line(const gh_graph* g, char* buf, int len) {
  return Guard([&] {
    const std::string s = g->weighted ? g->wg.StatsLine() : g->g.StatsLine();
    std::snprintf(buf, static_cast<size_t>(len), "%s", s.c_str());
    return static_cast<int>(s.size());
  });
}

const int64_t* gh_out_offsets(const gh_graph* g) {
  return g->weighted ? g->wg.out_offsets() : g->g.out_offsets();
}
const void* gh_out_dests(const gh_graph* g) {
  return g->weighted ? static_cast<const void*>(g->wg.out_dests())
                     : static_cast<const void*>(g->g.out_dests());
}
const int64_t* gh_in_offsets(const gh_graph* g) {
  return g->weighted ? g->wg.in_offsets() : g->g.in_offsets();
}
const void* gh_in_dests(const gh_graph* g) {
  return g->weighted ? static_cast<const void*>(g->wg.in_dests())
                     : static_cast<const void*>(g->g.in_dests());
}

int gh_strip_weights(const gh_graph* g, gh_graph** out) {
  return Guard([&] {
    *out = Wrap(StripWeights(Weighted(g)));
    return 0;
  });
}

int gh_relabel_by_degree(const gh_graph* g, gh_graph** out) {
  return Guard([&] {
    *out = Wrap(RelabelByDegree(Unweighted(g)));
    return 0;
  });
}

int gh_pack(const gh_graph* const* graphs, int count, gh_graph** out, int64_t* bounds) {
  return Guard([&] {
    if (count < 1 || !graphs) throw Error(-2, "no graphs to pack");
    std::vector<const Graph*> parts;
    for (int i = 0; i < count; i++) parts.push_back(&Unweighted(graphs[i]));
    std::vector<int64_t> b;
    *out = Wrap(PackGraphs(parts, &b));
    if (bounds) std::copy(b.begin(), b.end(), bounds);
    return 0;
  });
}

int gh_worth_relabelling(const gh_graph* g) {
  return Guard([&] { return WorthRelabelling(Unweighted(g)) ? 1 : 0; });
}

int gh_pick_sources(const gh_graph* g, int64_t given, int k, int32_t* out) {
  return Guard([&] {
    if (g->weighted) {
      SourcePicker<WGraph> sp(g->wg, static_cast<NodeID>(given));
      for (int i = 0; i < k; i++) out[i] = sp.PickNext();
    } else {
      SourcePicker<Graph> sp(g->g, static_cast<NodeID>(given));
      for (int i = 0; i < k; i++) out[i] = sp.PickNext();
    }
    return 0;
  });
}

int gh_verify_bfs(const gh_graph* g, int32_t source, const int32_t* parent) {
  return Guard([&] { return BFSVerifier(Unweighted(g), source, parent) ? 1 : 0; });
}

int gh_verify_sssp(const gh_graph* g, int32_t source, const int32_t* dist) {
  return Guard([&] { return SSSPVerifier(Weighted(g), source, dist) ? 1 : 0; });
}

int gh_verify_pr(const gh_graph* g, const float* scores, double tolerance) {
  return Guard([&] { return PRVerifier(Unweighted(g), scores, tolerance) ? 1 : 0; });
}

int gh_verify_cc(const gh_graph* g, const int32_t* comp) {
  return Guard([&] { return CCVerifier(Unweighted(g), comp) ? 1 : 0; });
}

int gh_verify_bc(const gh_graph* g, const int32_t* roots, int k, const float* scores) {
  return Guard([&] {
    const Graph& gr = Unweighted(g);
    const std::vector<float> oracle = SerialBrandes(gr, std::vector<NodeID>(roots, roots + k));
    bool ok = true;
    for (NodeID n = 0; n < gr.num_nodes(); n++)
      if (std::fabs(scores[n] - oracle[n]) > std::numeric_limits<float>::epsilon()) ok = false;
    return ok ? 1 : 0;
  });
}

int gh_verify_tc(const gh_graph* g, uint64_t count) {
  return Guard([&] { return TCVerifier(Unweighted(g), count) ? 1 : 0; });
}

const char* gh_last_error(void) { return g_error.c_str(); }

}  // extern "C"
