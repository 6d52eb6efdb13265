/* gaphost -- C-ABI of the host-side graph layer (libgaphost.so).
 *
 * The reference has no FFI; its host layer is header-only C++ that every
 * kernel binary compiles in.  This library is that layer for callers that
 * are not C++ (the Python tests and bench.py): graph generation/building,
 * .sg/.wsg I/O, root picking and the per-kernel verifiers.  The GPU kernels
 * live behind include/gapgpu.h; a graph built here is handed to them as
 * plain arrays (gh_out_offsets / gh_out_dests ...).
 *
 * Conventions: every call returns 0 (or a count / boolean >= 0) on success
 * and a negative status on failure -- the exit status the reference uses for
 * the same condition where it has one; gh_last_error() then describes it.
 * No call exits the process.
 */
#ifndef GAPHOST_H_
#define GAPHOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gh_graph gh_graph;

enum gh_kind { GH_KRONECKER = 0, GH_UNIFORM = 1, GH_GRID = 2 };

/* Generates and builds a symmetrised graph.
 * Replaces BuilderBase::MakeGraph's generator branch (src/builder.h:355-358),
 * Generator::GenerateEL (src/generator.h:157-168) with an explicit seed.
 * kind GH_GRID: `scale` is the grid side, `degree` is ignored. */
int gh_generate(int kind, int scale, int degree, int64_t seed, int weighted, gh_graph** out);

/* The vertex permutation of Generator::PermuteIDs (src/generator.h:93-103):
 * perm[i] = i shuffled by std::shuffle with an mt19937 seeded with `seed`.
 * The Kronecker generator maps every endpoint v to perm[v]; gg_graph_generate
 * takes this array. */
int gh_permutation(int64_t n, int64_t seed, int32_t* perm);

/* The generator's edge list before the build: Generator::GenerateEL
 * (src/generator.h:157-168) with an explicit seed, plus InsertWeights
 * (generator.h:170-187) when weighted -- the list gh_generate builds.  On
 * input *m is the capacity of u, v (and w when weighted); on output the
 * edge count.  With u NULL only the count is returned (for GH_KRONECKER and
 * GH_UNIFORM it is degree << scale). */
int gh_generate_edges(int kind, int scale, int degree, int64_t seed, int weighted, int64_t* m, int32_t* u,
                      int32_t* v, int32_t* w);

/* Builds a graph from an edge list (u[i] -> v[i], optional weights w[i]).
 * Replaces BuilderBase::MakeGraphFromEL + SquishGraph (src/builder.h:317-341, 172-186). */
int gh_build_edges(const int32_t* u, const int32_t* v, const int32_t* w, int64_t m,
                   int symmetrize, int weighted, gh_graph** out);

/* Reads / writes .sg (weighted = 0) or .wsg (weighted = 1).
 * Replaces Reader::ReadSerializedGraph (src/reader.h:254-307) and
 * WriterBase::WriteSerializedGraph (src/writer.h:39-69). */
int gh_read(const char* path, int weighted, gh_graph** out);
int gh_write(const gh_graph* g, const char* path);

void gh_free(gh_graph* g);

/* nodes, stored out-entries (num_edges_directed), directed, weighted. */
int gh_shape(const gh_graph* g, int64_t* nodes, int64_t* entries, int* directed, int* weighted);
/* "Graph has N nodes and M (un)directed edges for degree: D" (src/graph.h:223-230). */
int gh_stats_line(const gh_graph* g, char* buf, int len);

/* Borrowed views, valid until gh_free.  Dests are int32 (unweighted) or
 * {int32 v, int32 w} pairs (weighted).  For undirected graphs the in-views
 * alias the out-views. */
const int64_t* gh_out_offsets(const gh_graph* g);
const void* gh_out_dests(const gh_graph* g);
const int64_t* gh_in_offsets(const gh_graph* g);
const void* gh_in_dests(const gh_graph* g);

/* Unweighted copy of a weighted graph (same topology). */
int gh_strip_weights(const gh_graph* g, gh_graph** out);
/* Builder::RelabelByDegree (src/builder.h:367-402). */
int gh_relabel_by_degree(const gh_graph* g, gh_graph** out);
/* Disjoint union of `count` unweighted graphs (all undirected or all
 * directed): graph i's vertex v becomes bounds[i] + v; bounds[count + 1]
 * (nullable) receives the part boundaries, bounds[0] = 0.  No reference
 * counterpart: it feeds gg_bfs_packed. */
int gh_pack(const gh_graph* const* graphs, int count, gh_graph** out, int64_t* bounds);
/* tc.cc:75-91 WorthRelabelling heuristic: 1 / 0. */
int gh_worth_relabelling(const gh_graph* g);

/* First k roots of SourcePicker(g, given) (src/benchmark.h:46-68). */
int gh_pick_sources(const gh_graph* g, int64_t given, int k, int32_t* out);

/* Verifiers: 1 = PASS, 0 = FAIL, < 0 = error. */
int gh_verify_bfs(const gh_graph* g, int32_t source, const int32_t* parent);
int gh_verify_sssp(const gh_graph* g, int32_t source, const int32_t* dist);
int gh_verify_pr(const gh_graph* g, const float* scores, double tolerance);
int gh_verify_cc(const gh_graph* g, const int32_t* comp);
int gh_verify_bc(const gh_graph* g, const int32_t* roots, int k, const float* scores);
int gh_verify_tc(const gh_graph* g, uint64_t count);

const char* gh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GAPHOST_H_ */
