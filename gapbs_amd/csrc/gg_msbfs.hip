// Batched BFS entry point (traversal in gg_msbfs.h).
//
// Reference counterpart: none directly -- the reference runs one root per
// trial (src/benchmark.h:106, src/bfs.cc:256-259).  Each root's depths
// equal what DOBFS gives for it alone (bfs.cc:122-179).
#include "gg_msbfs.h"

using namespace gg;

extern "C" int gg_bfs_batch(gg_graph* g, const int32_t* roots, int k, int alpha, int32_t* depths,
                            gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "BFS needs an unweighted graph upload");
    GG_REQUIRE(k >= 1 && k <= 64 && roots, GG_EINVAL, "1 to 64 roots");
    GG_REQUIRE(alpha > 0, GG_EINVAL, "alpha must be positive");
    for (int i = 0; i < k; i++) GG_REQUIRE(roots[i] >= 0 && roots[i] < g->n, GG_EINVAL, "root out of range");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    hipStream_t st = g->stream;
    const dim3 blk(kBS);
    const dim3 grid_n(GridFor(n, kBS, g->cu_count * 16));
    uint16_t* depth16 = g->Scratch<uint16_t>("ms.depth16", n * 64);
    int32_t* droots = g->Scratch<int32_t>("ms.roots", 64);
    GG_HIP(hipMemcpyAsync(droots, roots, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    BeginTiming(g);
    const MsTraversal t = MsTraverse(g, droots, k, alpha, depth16, false);
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (depths) {
      int32_t* d32 = g->Scratch<int32_t>("ms.depth_out", n);
      for (int r = 0; r < k; r++) {
        MsDepthOut<<<grid_n, blk, 0, st>>>(depth16, n, r, d32);
        GG_HIP(hipMemcpyAsync(depths + int64_t(r) * n, d32, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
      }
      GG_HIP(hipStreamSynchronize(st));
    }
    if (stats) {
      unsigned long long* sums = g->Scratch<unsigned long long>("ms.reach", 2);
      MsZero2<<<1, 64, 0, st>>>(sums);
      MsReach<<<grid_n, blk, 0, st>>>(t.seen, g->out_off, n, sums);
      unsigned long long h[2];
      GG_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      stats->total_ms = wt.ms();
      stats->iterations = t.td + t.bu;
      stats->td_steps = t.td;
      stats->bu_steps = t.bu;
      stats->reached = static_cast<int64_t>(h[0]);
      stats->reached_edges = static_cast<int64_t>(h[1]);
      stats->edges_examined = t.examined;
      stats->launches = t.launches;
    }
    GG_HIP(hipGetLastError());
  });
}
