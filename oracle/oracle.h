/* CPU oracle for the kernel hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of the reference's kernels over plain CSR arrays
 * (int64 offsets[n+1], int32 destinations, {int32 v, int32 w} pairs when
 * weighted), used by tests/ as the parity checker and by bench.py as the CPU
 * baseline ("port").  Nothing in the product links or calls this library.
 *
 * Each function cites the reference code it restates.  How each one is
 * pinned (golden values, the reference built from source, closed forms,
 * brute force) is listed in the tests that exercise it.
 */
#ifndef ORACLE_H_
#define ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_bfs_stats {
  int64_t td_steps, bu_steps;
  int64_t edges_examined;  /* adjacency entries read under the reference schedule */
  int64_t vertices_examined; /* TD frontier sizes + BU unvisited vertices scanned */
  int64_t reached, reached_edges;
} orc_bfs_stats;

/* Per-step record of orc_dobfs (filled when steps != NULL, up to max_steps). */
typedef struct orc_bfs_step {
  int32_t bottom_up;
  int32_t pad;
  int64_t frontier;          /* TD: queue size; BU: vertices discovered in the previous step */
  int64_t vertices_examined; /* TD: frontier size; BU: unvisited vertices scanned */
  int64_t edges_examined;
  int64_t discovered;
} orc_bfs_step;

/* Serial level-synchronous BFS depths (textbook; src/bfs.cc:204-215). */
void orc_bfs_depth(int64_t n, const int64_t* off, const int32_t* dst, int32_t src, int32_t* depth);

/* OpenMP direction-optimizing BFS, restating DOBFS (src/bfs.cc:46-179).
 * Writes parent[n]; returns elapsed seconds.  ioff/idst: inverse CSR (the
 * out-CSR again for undirected graphs).  e_dir = num_edges_directed. */
double orc_dobfs(int64_t n, const int64_t* off, const int32_t* dst, const int64_t* ioff,
                 const int32_t* idst, int64_t e_dir, int32_t src, int alpha, int beta,
                 int32_t* parent, orc_bfs_stats* st);
/* Same, also recording each step into steps[0 .. max_steps) and returning
 * the step count through *nsteps. */
double orc_dobfs_trace(int64_t n, const int64_t* off, const int32_t* dst, const int64_t* ioff,
                       const int32_t* idst, int64_t e_dir, int32_t src, int alpha, int beta,
                       int32_t* parent, orc_bfs_stats* st, orc_bfs_step* steps, int max_steps,
                       int* nsteps);

/* Serial binary-heap Dijkstra (textbook; the reference verifier's oracle,
 * src/sssp.cc:164-194).  wdst: {v, w} pairs.  Unreached: INT32_MAX / 2. */
void orc_dijkstra(int64_t n, const int64_t* off, const int32_t* wdst, int32_t src, int32_t* dist);

/* OpenMP delta-stepping restating DeltaStep (src/sssp.cc:68-153): shared
 * frontier, thread-local bins, bucket fusion below 1000 entries.  Returns
 * elapsed seconds. */
double orc_delta_step(int64_t n, const int64_t* off, const int32_t* wdst, int64_t e_dir, int32_t src,
                      int32_t delta, int32_t* dist);

/* Jacobi pull PageRank restating PageRankPull (src/pr_spmv.cc:35-61) in
 * float (bit-identical to the reference: each vertex's sum is taken in
 * in-neighbour order) -- orc_pr_pull_f32 -- and the same iteration in
 * double with the reference's float constants widened -- orc_pr_pull_f64,
 * the FP64 oracle.  in_off/in_dst: inverse CSR; out_off: out-degrees.
 * Return the iterations run; *error receives the last L1 change. */
int orc_pr_pull_f32(int64_t n, const int64_t* in_off, const int32_t* in_dst, const int64_t* out_off,
                    int max_iters, double epsilon, float* scores, double* error);
int orc_pr_pull_f64(int64_t n, const int64_t* in_off, const int32_t* in_dst, const int64_t* out_off,
                    int max_iters, double epsilon, double* scores, double* error);

/* OpenMP Afforest restating Afforest/Link/Compress/SampleFrequentElement
 * (src/cc.cc:41-151).  ioff/idst: inverse CSR (read for directed graphs).
 * Returns elapsed seconds. */
double orc_afforest(int64_t n, const int64_t* off, const int32_t* dst, int directed,
                    const int64_t* ioff, const int32_t* idst, int neighbor_rounds, int32_t* comp);

/* Textbook components (serial union-find over every edge, union by
 * smaller id): label = smallest vertex id of the component. */
void orc_cc_min_labels(int64_t n, const int64_t* off, const int32_t* dst, int32_t* comp);

/* Certificate that dist holds the shortest-path distances from src in a
 * symmetric graph with positive weights (wdst interleaved (dest, weight),
 * or unit weights over dst when wdst is null; unreached vertices hold
 * `unreached`): dist[src] == 0, no edge out of a reached vertex relaxes or
 * leads to an unreached one, and every other reached vertex has a
 * neighbour on a shortest path.  Returns how many vertices break it. */
int64_t orc_distance_certificate(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* wdst,
                                 int32_t src, const int32_t* dist, int32_t unreached);

/* Brandes over roots[0 .. k), max-normalised, restating the reference
 * verifier's algorithm (src/bc.cc:169-220): double path counts, successors
 * from depths.  Each root is the serial algorithm; roots run one per thread
 * and their dependencies are added in root order, so the scores are those
 * of the serial loop over roots.  _f32: float dependencies and scores (the
 * reference's types, bc.cc:46-48); _f64: double throughout (FP64 oracle). */
void orc_brandes_f32(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* roots, int k,
                     float* scores);
void orc_brandes_f64(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* roots, int k,
                     double* scores);

/* OpenMP OrderedCount restatement (src/tc.cc:52-71); returns the count,
 * *seconds the elapsed time. */
uint64_t orc_ordered_count(int64_t n, const int64_t* off, const int32_t* dst, double* seconds);
uint64_t orc_ordered_count_sample(int64_t n, const int64_t* off, const int32_t* dst, int64_t stride,
                                  int64_t first, double* seconds);
/* Every triangle counted six times by sorted-list intersection over all
 * directed edges, divided by 6 (the reference verifier's method,
 * src/tc.cc:109-128). */
uint64_t orc_tc_intersect(int64_t n, const int64_t* off, const int32_t* dst);

/* Depths implied by a parent tree (-1 where unreached); returns -1 if the
 * parent pointers do not form a tree rooted at src. */
int orc_depth_from_parent(int64_t n, const int32_t* parent, int32_t src, int32_t* depth);

#ifdef __cplusplus
}
#endif

#endif /* ORACLE_H_ */
