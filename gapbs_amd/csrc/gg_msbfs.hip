// Batched BFS entry point (traversal in gg_msbfs.h).
//
// Reference counterpart: none directly -- the reference runs one root per
// trial (src/benchmark.h:106, src/bfs.cc:256-259).  Each root's depths
// equal what DOBFS gives for it alone (bfs.cc:122-179).
#include <algorithm>

#include "gg_msbfs.h"

namespace gg {
namespace {

// depth16[v][r] -> depth32[v] of root r (-1 where seen has no bit r: the
// slot is undefined there); with new_id, out is indexed by original ids.
__global__ __launch_bounds__(kBS) void MsDepthOut(const uint16_t* __restrict__ d16, const uint64_t* __restrict__ seen,
                                                  const int32_t* __restrict__ new_id, int64_t n, int r,
                                                  int32_t* __restrict__ out) {
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS) {
    const int64_t dv = new_id ? int64_t(new_id[v]) : v;
    const uint16_t d = (seen[dv] >> r) & 1 ? d16[dv * 64 + r] : kNoDepth;
    out[v] = d == kNoDepth ? -1 : d;
  }
}

}  // namespace
}  // namespace gg

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
    DeviceVertices(g, droots, k);
    BeginTiming(g);
    const MsTraversal t = MsTraverse(g, droots, nullptr, k, k == 64 ? ~0ull : (1ull << k) - 1, alpha, depth16, false);
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (depths) {
      int32_t* d32 = g->Scratch<int32_t>("ms.depth_out", n);
      for (int r = 0; r < k; r++) {
        MsDepthOut<<<grid_n, blk, 0, st>>>(depth16, t.seen, g->new_id, n, r, d32);
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

namespace gg {
namespace {

// Checks that every list entry lies in its vertex's part (every entry, so
// lists need not be sorted); sets *bad otherwise.  A wavefront per vertex
// range: each lane finds its vertex's part once, then the wavefront walks
// the lists of its 64 vertices with coalesced reads.
__global__ __launch_bounds__(kBS) void PkCheckParts(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                    int64_t n, const int64_t* __restrict__ part, int nparts,
                                                    int32_t* bad) {
  const int lane = LaneId();
  const int64_t wave = (blockIdx.x * int64_t(kBS) + threadIdx.x) / kWave;
  const int64_t waves = int64_t(gridDim.x) * kBS / kWave;
  bool wrong = false;
  for (int64_t v0 = wave * kWave; v0 < n; v0 += waves * kWave) {
    const int64_t v = v0 + lane;
    int64_t lo_id = 0, hi_id = 0, b = 0, e = 0;
    if (v < n) {
      int lo = 0, hi = nparts;  // part[lo] <= v < part[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (part[mid] <= v)
          lo = mid;
        else
          hi = mid;
      }
      lo_id = part[lo];
      hi_id = part[lo + 1];
      b = off[v];
      e = off[v + 1];
    }
    const int last = static_cast<int>(n - v0 < kWave ? n - v0 : kWave);
    for (int j = 0; j < last; j++) {
      const int64_t jb = __shfl(b, j), je = __shfl(e, j);
      const int64_t jlo = __shfl(lo_id, j), jhi = __shfl(hi_id, j);
      for (int64_t k = jb + lane; k < je; k += kWave) {
        const int32_t d = dst[k];
        wrong |= d < jlo || d >= jhi;
      }
    }
  }
  if (__ballot(wrong) && lane == 0) atomicOr(bad, 1);
}

// depth16[v][j] -> out[R * part[i] + j * |part i| + (v - part[i])] for the
// part i holding v, -1 where unreached (no bit j in seen[v]).
__global__ __launch_bounds__(kBS) void PkDepthOut(const uint16_t* __restrict__ d16,
                                                  const uint64_t* __restrict__ seen, int64_t n,
                                                  const int64_t* __restrict__ part, int nparts, int R,
                                                  int32_t* __restrict__ out) {
  const int64_t total = n * R;
  for (int64_t x = blockIdx.x * int64_t(kBS) + threadIdx.x; x < total; x += int64_t(gridDim.x) * kBS) {
    const int64_t v = x / R;
    const int j = static_cast<int>(x % R);
    int lo = 0, hi = nparts;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (part[mid] <= v)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t size = part[lo + 1] - part[lo];
    const uint16_t d = (seen[v] >> j) & 1 ? d16[v * 64 + j] : kNoDepth;
    out[R * part[lo] + j * size + (v - part[lo])] = d == kNoDepth ? -1 : d;
  }
}

}  // namespace
}  // namespace gg

extern "C" int gg_bfs_packed(gg_graph* g, const int64_t* part, int nparts, const int32_t* roots, int roots_per_part,
                             int alpha, int32_t* depths, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "BFS needs an unweighted graph upload");
    GG_REQUIRE(nparts >= 1 && part && roots, GG_EINVAL, "null part or root list");
    const int R = roots_per_part;
    GG_REQUIRE(R >= 1 && R <= 64, GG_EINVAL, "1 to 64 roots per part");
    GG_REQUIRE(alpha > 0, GG_EINVAL, "alpha must be positive");
    GG_REQUIRE(!g->new_id, GG_ENOTSUP, "packed traversal needs the uploaded vertex order (parts are id ranges)");
    GG_REQUIRE(part[0] == 0 && part[nparts] == g->n, GG_EINVAL, "parts must cover the vertices");
    for (int i = 0; i < nparts; i++) {
      GG_REQUIRE(part[i] < part[i + 1], GG_EINVAL, "parts must be non-empty and increasing");
      for (int j = 0; j < R; j++) {
        const int32_t r = roots[int64_t(i) * R + j];
        GG_REQUIRE(r >= part[i] && r < part[i + 1], GG_EINVAL, "root outside its part");
      }
    }
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    const int64_t count = int64_t(nparts) * R;
    GG_REQUIRE(count < (int64_t(1) << 30), GG_EINVAL, "too many roots");
    hipStream_t st = g->stream;
    const dim3 blk(kBS);
    const dim3 grid_n(GridFor(n, kBS, g->cu_count * 16));
    int64_t* dpart = g->Scratch<int64_t>("pk.part", nparts + 1);
    int32_t* droots = g->Scratch<int32_t>("pk.roots", count);
    int32_t* dbits = g->Scratch<int32_t>("pk.bits", count);
    int32_t* bad = g->Scratch<int32_t>("pk.bad", 1);
    std::vector<int32_t> bits(count);
    for (int64_t x = 0; x < count; x++) bits[x] = static_cast<int32_t>(x % R);
    GG_HIP(hipMemcpyAsync(dpart, part, sizeof(int64_t) * (nparts + 1), hipMemcpyHostToDevice, st));
    GG_HIP(hipMemcpyAsync(droots, roots, sizeof(int32_t) * count, hipMemcpyHostToDevice, st));
    GG_HIP(hipMemcpyAsync(dbits, bits.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, st));
    // The lists are checked against the parts once per handle and part
    // list (the graph does not change between calls).
    if (g->pk_parts.size() != static_cast<size_t>(nparts) + 1 ||
        !std::equal(g->pk_parts.begin(), g->pk_parts.end(), part)) {
      g->pk_parts.clear();
      GG_HIP(hipMemsetAsync(bad, 0, sizeof(int32_t), st));
      PkCheckParts<<<grid_n, blk, 0, st>>>(g->out_off, g->out_dst, n, dpart, nparts, bad);
      if (g->directed) PkCheckParts<<<grid_n, blk, 0, st>>>(g->in_off, g->in_dst, n, dpart, nparts, bad);
      int32_t hbad = 0;
      GG_HIP(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      GG_REQUIRE(hbad == 0, GG_EINVAL, "an edge joins two parts");
      g->pk_parts.assign(part, part + nparts + 1);
    }
    uint16_t* depth16 = g->Scratch<uint16_t>("ms.depth16", n * 64);
    const uint64_t full = R == 64 ? ~0ull : (1ull << R) - 1;
    BeginTiming(g);
    const MsTraversal t = MsTraverse(g, droots, dbits, static_cast<int>(count), full, alpha, depth16, false);
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (depths) {
      int32_t* d32 = g->Scratch<int32_t>("pk.depth_out", n * R);
      PkDepthOut<<<dim3(GridFor(n * R, kBS, g->cu_count * 16)), blk, 0, st>>>(depth16, t.seen, n, dpart, nparts, R,
                                                                                d32);
      GG_HIP(hipMemcpyAsync(depths, d32, sizeof(int32_t) * n * R, hipMemcpyDeviceToHost, st));
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
