// Ordered triangle counting on gfx950.
//
// Reference: OrderedCount (src/tc.cc:52-71).  Each triangle u > v > w is
// counted once, at its largest vertex u: for every neighbour v < u of u,
// the neighbours w of u below v that are also neighbours of v, found by a
// merge of the two sorted lists with an early exit once w reaches v.  The
// graph must be undirected with sorted, duplicate-free lists (the builder
// guarantees it, tc.cc:31-35).  Whether to relabel by degree first
// (Hybrid / WorthRelabelling, tc.cc:75-100) is the caller's choice; the
// relabelled graph is counted the same way.
//
// GPU design: the forward edges u -> v (v < u, a prefix of u's sorted list,
// found by binary search) are spread over the grid with gg_expand.h; each
// lane merges one (u, v) pair.  Counts are summed per block and added to 64
// striped counters.
#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kStripes = 64;

struct TcVisit {
  const int64_t* off;
  const int32_t* dst;
  unsigned long long count = 0, work = 0;
  __device__ __forceinline__ void operator()(int32_t u, int32_t, int64_t e, bool valid) {
    if (!valid) return;
    const int32_t v = dst[e];
    const int64_t ue = off[u + 1];
    int64_t it = off[v];
    const int64_t ve = off[v + 1];
    for (int64_t iu = off[u]; iu < ue; iu++) {
      const int32_t w = dst[iu];
      if (w > v) break;
      while (it < ve && dst[it] < w) it++;
      work++;
      if (it < ve && dst[it] == w) count++;
    }
  }
};

// Index of the first entry of the sorted range [b, e) that is >= x.
__device__ __forceinline__ int64_t LowerBound(const int32_t* dst, int64_t b, int64_t e, int32_t x) {
  while (b < e) {
    const int64_t m = b + ((e - b) >> 1);
    if (dst[m] < x)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

__global__ __launch_bounds__(kBS) void TcCount(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                               int64_t n, unsigned long long* counts, EdgeChunk* chunks,
                                               int32_t* nchunks) {
  __shared__ ExpandShared sh;
  __shared__ unsigned long long s_red[kBS / kWave];
  TcVisit visit{off, dst};
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < n; tile += int64_t(gridDim.x) * kBS) {
    const int64_t u = tile + threadIdx.x;
    int64_t beg = 0, end = 0;
    if (u < n) {
      beg = off[u];
      end = LowerBound(dst, beg, off[u + 1], static_cast<int32_t>(u));  // forward edges v < u
    }
    ExpandTile(sh, static_cast<int32_t>(u), 0, beg, end, visit, chunks, nchunks, kChunkEdges);
  }
  const unsigned long long c = BlockSum<kBS>(visit.count, s_red);
  const unsigned long long w = BlockSum<kBS>(visit.work, s_red);
  if (threadIdx.x == 0) {
    if (c) atomicAdd(&counts[2 * (blockIdx.x & (kStripes - 1))], c);
    if (w) atomicAdd(&counts[2 * (blockIdx.x & (kStripes - 1)) + 1], w);
  }
}

__global__ __launch_bounds__(kBS) void TcCountChunks(const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ dst,
                                                     unsigned long long* counts, const EdgeChunk* chunks,
                                                     const int32_t* nchunks) {
  __shared__ unsigned long long s_red[kBS / kWave];
  TcVisit visit{off, dst};
  ExpandChunks(chunks, *nchunks, visit);
  const unsigned long long c = BlockSum<kBS>(visit.count, s_red);
  const unsigned long long w = BlockSum<kBS>(visit.work, s_red);
  if (threadIdx.x == 0) {
    if (c) atomicAdd(&counts[2 * (blockIdx.x & (kStripes - 1))], c);
    if (w) atomicAdd(&counts[2 * (blockIdx.x & (kStripes - 1)) + 1], w);
  }
}

__global__ void TcClear(unsigned long long* counts, int32_t* nchunks) {
  for (int i = threadIdx.x; i < 2 * kStripes; i += blockDim.x) counts[i] = 0;
  if (threadIdx.x == 0) *nchunks = 0;
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_tc(gg_graph* g, uint64_t* count, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "TC needs an unweighted graph upload");
    GG_REQUIRE(!g->directed, GG_ENOTSUP, "TC needs an undirected graph");
    GG_REQUIRE(count != nullptr, GG_EINVAL, "null count output");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    hipStream_t st = g->stream;
    unsigned long long* counts = g->Scratch<unsigned long long>("tc.counts", 2 * kStripes);
    int32_t* nchunks = g->Scratch<int32_t>("tc.nchunks", 1);
    const bool use_chunks = g->max_degree >= kChunkEdges;
    EdgeChunk* chunks = use_chunks ? g->Scratch<EdgeChunk>("tc.chunks", MaxChunks(n, g->out_entries)) : nullptr;
    const dim3 blk(kBS);
    int64_t launches = 0;
    BeginTiming(g);
    Run(g, "tc_clear", 0, TcClear, dim3(1), blk, counts, nchunks);
    // One pass over offsets and lists, plus the list reads of the merges
    // (credited from the work counter below).
    const int idx = Run(g, "tc_count", 8.0 * (n + 1) + 4.0 * g->out_entries, TcCount,
                        dim3(GridFor(n, kBS, g->cu_count * 8)), blk, static_cast<const int64_t*>(g->out_off),
                        static_cast<const int32_t*>(g->out_dst), n, counts,
                        use_chunks ? chunks : static_cast<EdgeChunk*>(nullptr), nchunks);
    launches += 2;
    if (use_chunks) {
      Run(g, "tc_count_chunks", 0, TcCountChunks, dim3(g->cu_count * 4), blk,
          static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst), counts,
          static_cast<const EdgeChunk*>(chunks), static_cast<const int32_t*>(nchunks));
      launches++;
    }
    GG_HIP(hipGetLastError());
    unsigned long long* h = g->Pinned<unsigned long long>(2 * kStripes);
    GG_HIP(hipMemcpyAsync(h, counts, sizeof(unsigned long long) * 2 * kStripes, hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    unsigned long long total = 0, work = 0;
    for (int i = 0; i < kStripes; i++) {
      total += h[2 * i];
      work += h[2 * i + 1];
    }
    CreditBytes(g, idx, 4.0 * static_cast<double>(work));  // merge reads
    FinishTiming(g, stats, wt);
    *count = total;
    if (stats) {
      stats->total_ms = wt.ms();
      stats->iterations = 1;
      stats->edges_examined = static_cast<int64_t>(work);
      stats->launches = launches;
    }
  });
}
