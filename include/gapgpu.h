/* gapgpu -- C-ABI of the MI355X (gfx950) kernel library (libgapgpu.so).
 *
 * The reference has no FFI, plugin registry or operator table.  Its kernel
 * boundary is the functor the harness calls once per trial,
 *   Result kernel(const Graph&)            (src/benchmark.h:98-106)
 * bound in each driver's main() to one kernel function.  Each entry point
 * below replaces one of those kernel functions; a driver wraps it in the
 * same lambda, so BenchmarkKernel, the analysis printers and the verifiers
 * stay as they are.
 *
 * Conventions
 *  - Return 0 on success, a negative GG_E* status on failure; never exit or
 *    abort.  gg_last_error() (thread-local) describes the last failure.
 *  - Graph arrays are the .sg/.wsg layout: int64 offsets[n+1] and int32
 *    destinations (weighted: {int32 v, int32 w} pairs).  gg_graph_upload
 *    copies them into HBM once; the handle owns that device memory.
 *  - Output buffers are caller-owned host memory sized as documented.
 *    Calls are synchronous: they return when the results are in host memory.
 *  - Each handle has its own HIP stream; use one handle from one host thread
 *    at a time (the reference's harness is single-threaded).
 *  - gg_stats (nullable) receives device time measured with HIP events on
 *    that stream, and per-kernel counters.
 */
#ifndef GAPGPU_H_
#define GAPGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GG_OK = 0,
  GG_EINVAL = -22,   /* bad argument (root out of range, wrong graph kind ...) */
  GG_ENODEV = -19,   /* no usable GPU / gg_init not called */
  GG_ENOMEM = -12,   /* device allocation failed or exceeds GAPGPU_MEM_CAP_BYTES */
  GG_EHIP = -5,      /* any other HIP runtime error */
  GG_ENOTSUP = -95   /* unsupported graph shape for this kernel */
};

typedef struct gg_graph gg_graph;

typedef struct gg_stats {
  double device_ms;       /* HIP-event time of the kernel sequence on the handle's stream */
  double total_ms;        /* host wall time of the call, result copies included */
  int64_t iterations;     /* levels (BFS/BC), buckets (SSSP), iterations (PR), passes (CC) */
  int64_t td_steps;       /* BFS: top-down levels */
  int64_t bu_steps;       /* BFS: bottom-up levels */
  int64_t reached;        /* vertices reached (BFS/SSSP/BC, summed over roots) */
  int64_t reached_edges;  /* sum of out-degrees of reached vertices (summed over roots) */
  int64_t edges_examined; /* adjacency entries read by the kernels, when counted */
  int64_t vertices_examined; /* BFS: frontier vertices expanded + unvisited vertices scanned */
  double residual;        /* PR: final L1 change (FP64) */
  int64_t launches;       /* kernel launches issued */
} gg_stats;

/* Selects and initialises a HIP device for the calling process. */
int gg_init(int device);

/* Copies a CSR graph into HBM.  weighted: dests are {v, w} pairs (2 int32 per
 * entry).  directed: the inverse CSR (in_offsets / in_dests) is required;
 * for undirected graphs pass NULL and the out-arrays serve both directions
 * (src/graph.h:138-143). */
int gg_graph_upload(int64_t num_nodes, const int64_t* out_offsets, const int32_t* out_dests,
                    int weighted, int directed, const int64_t* in_offsets, const int32_t* in_dests,
                    gg_graph** out);
int gg_graph_free(gg_graph* g);
/* Device bytes held by the handle (graph + cached kernel workspaces). */
int64_t gg_graph_device_bytes(const gg_graph* g);

/* Direction-optimizing BFS from `source`.
 * Replaces DOBFS(const Graph&, NodeID source, bool logging, int alpha=15,
 * int beta=18) (src/bfs.cc:122-179) with the same alpha/beta switch.
 * parent[n]: parent[source] = source, -1 if unreached.
 * depth[n]: BFS level, -1 if unreached.
 * Either output may be NULL: the result then stays in device memory only
 * (used to time the kernels without the result copy). */
int gg_bfs(gg_graph* g, int32_t source, int alpha, int beta, int32_t* parent, int32_t* depth,
           gg_stats* stats);

/* Batched BFS: k (1..64) roots traversed together with 64-bit per-vertex
 * masks.  Replaces k sequential DOBFS trials (src/benchmark.h:106,
 * src/bfs.cc:256-259); each root's depths equal its single-root BFS depths.
 * depths[k * n] (nullable: kept on the device), root-major: depths[r * n +
 * v] is v's level from roots[r], -1 if unreached.  Repeated roots are
 * allowed.  alpha: go bottom-up when the frontier's out-edges exceed the
 * unexplored edges / alpha (as DOBFS).  stats->reached counts reached
 * (root, vertex) pairs; stats->reached_edges sums their out-degrees. */
int gg_bfs_batch(gg_graph* g, const int32_t* roots, int k, int alpha, int32_t* depths, gg_stats* stats);

/* Packed batched BFS: the upload is the disjoint union of nparts graphs,
 * part i holding vertices [part[i], part[i+1]) (part[0] = 0, part[nparts] =
 * n); no edge may join two parts (checked: GG_EINVAL).  roots[i * R + j]
 * (R = roots_per_part, 1..64) is root j of part i and must lie in part i.
 * All nparts * R roots run in one traversal: root j of every part takes
 * visit-mask bit j (parts never meet, so bits are shared safely).  No
 * reference counterpart: the reference runs one graph and one root per
 * trial (src/benchmark.h:106); each root's depths equal its single-root BFS
 * depths on its own graph.  depths[R * n] (nullable), part-major: part i's
 * block starts at R * part[i] and holds R rows of |part i| depths (-1 if
 * unreached), row j for root j. */
int gg_bfs_packed(gg_graph* g, const int64_t* part, int nparts, const int32_t* roots, int roots_per_part,
                  int alpha, int32_t* depths, gg_stats* stats);

/* Delta-stepping single-source shortest paths from `source` with bucket
 * width `delta` on a weighted upload.
 * Replaces DeltaStep(const WGraph&, NodeID source, WeightT delta, bool
 * logging) (src/sssp.cc:87-153).  dist[n] (nullable: kept on the device):
 * shortest distance, 1073741823 (INT32_MAX / 2) if unreached. */
int gg_sssp(gg_graph* g, int32_t source, int32_t delta, int32_t* dist, gg_stats* stats);

/* Batched delta-stepping: k (1..64) roots searched in one traversal that
 * shares the buckets (width delta) and the edge reads between roots; the
 * bucket loop runs on the device (its state in HBM, rounds replayed from a
 * captured hipGraph).  Replaces k sequential DeltaStep trials
 * (src/benchmark.h:106 around src/sssp.cc:87-153); each root's distances
 * equal its gg_sssp distances bit for bit.  dist[k * n] (nullable: kept on
 * the device), root-major: dist[r * n + v] is v's distance from roots[r],
 * 1073741823 if unreached.  Repeated roots are allowed.  Device memory
 * beyond the graph: 4 G + 40 bytes per vertex, G = the smallest power of two
 * >= k (plus 4 k per vertex for the copy-out).  stats->reached counts
 * reached (root, vertex) pairs, stats->reached_edges sums their out-degrees,
 * stats->iterations counts buckets, stats->td_steps relax launches and
 * stats->edges_examined (edge, root) relaxations. */
int gg_sssp_batch(gg_graph* g, const int32_t* roots, int k, int32_t delta, int32_t* dist, gg_stats* stats);

/* Pull PageRank with damping 0.85.
 * Replaces PageRankPull(const Graph&, int max_iters, double epsilon, bool
 * logging) (src/pr_spmv.cc:35-61): Jacobi iterations until the L1 change
 * (accumulated in double) drops below epsilon or max_iters is reached.
 * mode GG_PR_FP32: float scores with compensated per-vertex sums;
 * mode GG_PR_FP64: double scores and sums.  mode GG_PR_GS: FP32 with one
 * contribution array updated in place as scores change, the Gauss-Seidel
 * sweep of PageRankPullGS (src/pr.cc:34-61, the reference's `pr` binary);
 * like the reference's, its result depends on the schedule and is checked
 * by the convergence rule of PRVerifier (pr.cc:65-94).  scores[n]
 * (nullable) receives float scores.  stats->iterations and stats->residual
 * report the iterations run and the last L1 change. */
enum { GG_PR_FP32 = 0, GG_PR_FP64 = 1, GG_PR_GS = 2 };
int gg_pagerank(gg_graph* g, int max_iters, double epsilon, int mode, float* scores, gg_stats* stats);

/* The source-segment layout gg_pagerank gathers through in a Jacobi mode
 * (GG_PR_FP32 or GG_PR_FP64) on this handle, built here if it is not yet
 * (all 0 when the mode runs through propagation blocks instead).
 * No reference counterpart (the reference pulls every list whole); for
 * measurement reports.  info[0] LDS segments, info[1] L2 segments, info[2]
 * (destination, sub-list) pairs, info[3] in-entries held by the segments,
 * info[4] the source id below which entries go by segment, info[5] the
 * layout's build time in microseconds; all 0 when the gather runs without
 * segments. */
int gg_pagerank_layout(gg_graph* g, int mode, int64_t* info);

/* The propagation-block layout a Jacobi gg_pagerank (mode GG_PR_FP32 or
 * GG_PR_FP64) runs through on this handle when gg_tuning.pr_blocks_log2
 * selects it (by default on graphs of 2^20 to 2^27 vertices), built here if
 * it is not yet; it replaces the source segments, and one layout serves
 * both precisions.  No reference counterpart (the reference pulls every
 * list whole, pr_spmv.cc:47-53); for measurement reports.  info[0] log2 of
 * the tile / bin width (0: no blocks on this handle), info[1] tiles (=
 * bins), info[2] in-entries after padding each (tile, bin) block to a
 * multiple of 8, info[3] chunks of the write phase, info[4] chunks of the
 * sum phase, info[5] the build time in microseconds. */
int gg_pagerank_blocks(gg_graph* g, int mode, int64_t* info);

/* Afforest connected components.
 * Replaces Afforest(const Graph&, bool logging, int32_t neighbor_rounds=2)
 * (src/cc.cc:95-151).  comp[n] (nullable): component label = the smallest
 * vertex id in the component, exactly the reference's labels. */
int gg_cc(gg_graph* g, int neighbor_rounds, int32_t* comp, gg_stats* stats);

/* Allocation-free Afforest.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * result as gg_cc (labels bit-identical), but every device buffer comes from
 * a fixed 64 MiB array that each GPU holds from gg_init until the process
 * ends: the call makes no hipMalloc, hipFree, pinned or library allocation.
 * The graph needs 4 n bytes (8 n more when degree-ordered) plus 24 bytes per
 * 4096-entry chunk of its hubs' lists and a few KB to fit; a larger graph is
 * GG_EINVAL, with the bytes it needs in gg_last_error() (use gg_cc).  All
 * calls on one GPU share the array and a call leaves its contents behind:
 * calls that overlap on one GPU (two host threads, two handles) race. */
int gg_cc_noalloc(gg_graph* g, int neighbor_rounds, int32_t* comp, gg_stats* stats);

/* Shiloach-Vishkin connected components, the reference's second CC binary.
 * Replaces ShiloachVishkin(const Graph&) (src/cc_sv.cc:49-83): hook every
 * out-edge's higher root under the lower label (min-max swap, so directed
 * graphs get their weak components), compress, repeat until no hook.  The
 * labels equal gg_cc's (the smallest vertex id of each component);
 * stats->iterations is the reference's iteration count. */
int gg_cc_sv(gg_graph* g, int32_t* comp, gg_stats* stats);

/* Brandes betweenness centrality from roots[0 .. k).
 * Replaces Brandes(const Graph&, SourcePicker<Graph>&, NodeID num_iters,
 * bool logging) (src/bc.cc:95-150): the caller passes the roots the
 * SourcePicker would draw.  Path counts are FP64.  mode GG_BC_FP32: the
 * reference's float arithmetic (dependencies accumulated in successor order,
 * float scores), which its verifier requires to FLT_EPSILON;
 * GG_BC_FP32_KAHAN: float dependencies and scores with Kahan-compensated
 * per-vertex dependency sums and score sums; GG_BC_FP64: double
 * dependencies and scores.
 * scores[n] (nullable) receives float scores divided by the largest. */
enum { GG_BC_FP32 = 0, GG_BC_FP64 = 1, GG_BC_FP32_KAHAN = 2 };
int gg_bc(gg_graph* g, const int32_t* roots, int k, int mode, float* scores, gg_stats* stats);

/* Batched Brandes: the k (1..64) roots traversed together (64-bit visit
 * masks per vertex, path counts and dependencies stored per (vertex, root)).
 * Replaces the reference's per-root loop in Brandes (src/bc.cc:108-140):
 * the result equals gg_bc over the same roots -- bit for bit in
 * GG_BC_FP32 (each root's dependencies are folded in successor order with
 * the reference's arithmetic and added to the scores in root order), to
 * FP64 rounding in GG_BC_FP64.  GG_BC_FP32_KAHAN returns GG_ENOTSUP.
 * Device memory beyond the graph: about (8 x levels + 512 + 256) bytes per
 * vertex in GG_BC_FP32 (1024 + 8 x levels in GG_BC_FP64).
 * scores[n] (nullable) receives float scores divided by the largest. */
int gg_bc_batch(gg_graph* g, const int32_t* roots, int k, int mode, float* scores, gg_stats* stats);

/* Ordered triangle count of the uploaded undirected graph.
 * Replaces OrderedCount(const Graph&) (src/tc.cc:52-71); the Hybrid
 * relabel-by-degree decision (tc.cc:75-100) is made by the caller, which
 * uploads the relabelled graph when it applies.  stats->edges_examined is
 * the number of (set, list-entry) probes. */
int gg_tc(gg_graph* g, uint64_t* count, gg_stats* stats);

/* Relabelling by degree on the device.  Replaces
 * BuilderBase::RelabelByDegree (src/builder.h:367-402), which TC's Hybrid
 * runs inside its timed trial (src/tc.cc:97): *out receives a new handle
 * (undirected, unweighted) whose vertices are numbered by descending
 * (degree, old id) and whose lists are sorted -- byte-identical to the
 * reference's result.  Directed or weighted uploads: GG_ENOTSUP ("Cannot
 * relabel directed graph", as the reference).  stats->device_ms is the
 * relabel's device time.  Free *out with gg_graph_free. */
int gg_relabel_by_degree(gg_graph* g, gg_graph** out, gg_stats* stats);

/* Builds a graph on the device from an edge list (u[i] -> v[i], optional
 * weights w[i], m edges in host memory).  Replaces the CSR half of the
 * reference's builder: BuilderBase::MakeGraphFromEL / MakeCSR / SquishCSR /
 * SquishGraph (src/builder.h:300-341, 141-186), with the vertex count taken
 * as the largest endpoint + 1 (FindMaxNodeID, builder.h:69-78).  The handle
 * holds exactly the CSR gh_build_edges builds from the same arguments:
 * symmetrize != 0 stores both directions of every edge; otherwise the graph
 * is directed with out- and in-lists; weighted != 0 stores {v, w} pairs
 * (weights 1 when w is NULL); lists are sorted by (v, w), duplicates keep the
 * smallest weight, self-loops are dropped.  *num_nodes (nullable) receives
 * n.  GG_EINVAL for a negative id, GG_ENOTSUP when the lists hold 2^32
 * entries or more or 2 * bits(n - 1) + bits(weight range) exceeds 64.
 * stats->device_ms covers the build after the edge list reached HBM. */
int gg_graph_build(const int32_t* u, const int32_t* v, const int32_t* w, int64_t m, int symmetrize,
                   int weighted, gg_graph** out, int64_t* num_nodes, gg_stats* stats);
/* Generates and builds on the device the symmetrised graph gh_generate
 * builds: kind 0 Kronecker (R-MAT), 1 uniform; m = degree << scale edges;
 * weighted: weights uniform in [1, 255].  Replaces Generator::GenerateEL and
 * InsertWeights (src/generator.h:105-187) -- each 2^18-edge block's mt19937
 * stream seeded with seed + block, one wavefront per block -- followed by
 * gg_graph_build.  perm (Kronecker only, 2^scale entries) is the vertex
 * permutation that gh_permutation returns: Generator::PermuteIDs
 * (generator.h:93-103), a sequential std::shuffle left on the host.
 * stats->device_ms covers generation and build. */
int gg_graph_generate(int kind, int scale, int degree, int64_t seed, int weighted, const int32_t* perm,
                      gg_graph** out, int64_t* num_nodes, gg_stats* stats);
/* Reads a .sg (weighted = 0) or .wsg (weighted = 1) file straight into a
 * new handle, streaming it through two pinned staging buffers so that file
 * reads overlap the copies to HBM; the host never holds the graph.
 * Replaces Reader::ReadSerializedGraph (src/reader.h:254-307) followed by
 * gg_graph_upload.  The suffix rules and messages are the reference's
 * (reader.h:254-266); an unreadable, short or inconsistent file is
 * GG_EINVAL.  *num_nodes (nullable) receives n. */
int gg_graph_read(const char* path, int weighted, gg_graph** out, int64_t* num_nodes, gg_stats* stats);
/* Copies a handle's inverse CSR (in-lists; the out-lists for undirected
 * graphs) back to host memory, like gg_graph_copy_out. */
int gg_graph_copy_in(const gg_graph* g, int64_t* offsets, int32_t* dests);

/* Copies a handle's CSR back to host memory (offsets[n + 1] and dests, each
 * nullable): the layout gg_graph_upload takes.  A degree-ordered handle
 * (below) returns its lists in device ids. */
int gg_graph_copy_out(const gg_graph* g, int64_t* offsets, int32_t* dests);

/* Degree-ordered device layout (no reference counterpart: the reference
 * keeps the builder's ids for every kernel but TC).  Renumbers the handle's
 * vertices in place by descending (degree, id) -- the order of
 * gg_relabel_by_degree -- so the lists, and every per-vertex array the
 * kernels gather through them, keep hubs together.  The handle still speaks
 * uploaded ids: roots and sources are mapped on the way in, and every
 * per-vertex result (parents, depths, distances, scores, component labels)
 * comes back indexed by, and holding, uploaded ids.  Unweighted lists are
 * sorted under the new ids; weighted lists keep their uploaded order.
 * Exactness: BFS depths, SSSP distances, component labels (the smallest
 * uploaded id per component) and triangle counts are unchanged; BFS parents
 * remain a valid BFS tree; PageRank and BC sum in the new list order
 * (within tolerance of the uploaded order; bit-exact with the uploaded-order
 * result of the relabelled graph).  gg_bfs_packed needs the uploaded order
 * (GG_ENOTSUP).  Undirected graphs only (GG_ENOTSUP otherwise); calling it
 * twice is GG_EINVAL.  stats->device_ms is the renumbering's device time. */
int gg_graph_order_by_degree(gg_graph* g, gg_stats* stats);

/* new_id[v] = device id of uploaded vertex v (the identity unless the
 * handle is degree-ordered). */
int gg_graph_vertex_order(gg_graph* g, int32_t* new_id);

/* Per-graph tuning (no reference counterpart; the reference's knobs are the
 * -a/-b style alpha/beta defaults in bfs.cc:122-123 and sssp's -d delta).
 * gg_autotune times candidate settings of the selected kernels on this
 * graph from the given roots (k >= 1; SourcePicker roots are the natural
 * choice) and keeps the fastest in the handle:
 *   GG_TUNE_BFS     gg_bfs alpha x beta over {2..240} x {2..162}
 *   GG_TUNE_BATCH   gg_bfs_batch / gg_bfs_packed alpha over {1..1000}, then
 *                   on a graph with propagation blocks the first bottom-up
 *                   level with and without them (bfs_blocks_pct 50 or -1)
 *   GG_TUNE_SSSP    gg_sssp delta over powers of two 1..4096 (weighted
 *                   uploads only; slow candidates are cut off early)
 *   GG_TUNE_PR      on a graph whose Jacobi PageRank runs through
 *                   propagation blocks at the default width: that width or
 *                   a neighbour (pr_blocks_log2), by 3-iteration runs;
 *                   otherwise the in-degree bin cut-offs (thread /
 *                   wavefront / block per vertex) and the gather launch
 *                   width, by coordinate descent over 3-iteration runs
 * alpha, beta and delta are arguments of their entry points: callers read
 * them back (gg_tuning_get) and pass them; the PageRank settings apply to
 * every later gg_pagerank on the handle.  gg_tuning_set installs a saved
 * tuning (GG_EINVAL for values out of range). */
typedef struct gg_tuning {
  int32_t bfs_alpha, bfs_beta;
  int32_t batch_alpha;
  int32_t sssp_delta;
  int32_t pr_small_max;      /* in-degree below this: one thread per vertex */
  int32_t pr_medium_max;     /* below this: one wavefront; else one block */
  int32_t pr_blocks_per_cu;  /* gather launch width */
  int32_t tuned;             /* GG_TUNE_* bits tuned so far */
  double best_ms[4];         /* device ms of the chosen setting: bfs, batch, sssp, pr */
  /* Jacobi PageRank source segments, [0] FP32 and [1] FP64.
   * pr_segment_log2: width of the segments gathered through L2; 0 picks the
   * default (2 MB of contributions), -1 turns segments off altogether (one
   * pull per vertex), 4..30 sets log2 of the width in vertices.
   * pr_lds_log2: segments staged in LDS, which come first: 0 or -1 none
   * (the default), 4..15 (FP32) or 4..14 (FP64) log2 of their width. */
  int32_t pr_segment_log2[2];
  int32_t pr_lds_log2[2];
  /* Default widths keep the leading segments of each kind whose in-entries
   * are at least this many per (destination, sub-list) pair; 0 picks the
   * default (3), else 1..64. */
  int32_t pr_seg_min_density;
  /* Batched BFS (and batched BC's traversal): the first bottom-up level
   * after a top-down one reads every in-entry once through the propagation
   * blocks (the layout FP32 Jacobi PageRank uses, see pr_blocks_log2) when
   * more than this percentage of the in-edges belong to vertices some root
   * has not reached; 0 picks the default (50, for batches of 32 roots or
   * more; GG_TUNE_BATCH sets 50 or -1 by timing both), -1 never, 1..100
   * (whatever the number of roots). */
  int32_t bfs_blocks_pct;
  /* Jacobi PageRank propagation blocks (both precisions): log2 of the
   * tile / bin width in vertices.  0 picks the default: on graphs of 2^20
   * to 2^27 vertices whose pr_segment_log2 for the precision is left at 0,
   * width 2^(ceil(log2 n) - 12) clamped to 2^10..2^14 (about 4096 tiles),
   * else no blocks; -1 never; 4..14 always (the graph then needs at most
   * 8192 tiles). */
  int32_t pr_blocks_log2;
} gg_tuning;
enum { GG_TUNE_BFS = 1, GG_TUNE_BATCH = 2, GG_TUNE_SSSP = 4, GG_TUNE_PR = 8 };
int gg_autotune(gg_graph* g, int what, const int32_t* roots, int k, gg_tuning* out);
int gg_tuning_get(const gg_graph* g, gg_tuning* out);
int gg_tuning_set(gg_graph* g, const gg_tuning* in);

/* Per-kernel measurement (no reference counterpart; used by bench.py).
 * While enabled, every kernel launch on the handle is bracketed by HIP
 * events on the handle's stream and credited with its algorithmic HBM bytes
 * (the compulsory-traffic model of each kernel).  gg_profile_get returns the
 * idx-th kernel's totals, or GG_EINVAL past the last one. */
int gg_profile_enable(gg_graph* g, int on);
int gg_profile_reset(gg_graph* g);
int gg_profile_get(gg_graph* g, int idx, char* name, int name_len, int64_t* launches,
                   double* total_ms, double* alg_bytes);

const char* gg_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GAPGPU_H_ */
