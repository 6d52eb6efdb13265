// Relabelling by degree on the device (gg_relabel_by_degree).
//
// Reference: BuilderBase::RelabelByDegree (src/builder.h:367-402), which
// TC's Hybrid runs inside its timed trial (src/tc.cc:97): vertices get new
// ids in the order of descending (degree, old id) -- std::sort with
// std::greater over (degree, id) pairs, so equal degrees keep the larger
// old id first -- every list is rewritten through the new ids and sorted.
// The result here is byte-identical to that.
//
// Device steps:
//   1. key = ~((degree << 32) | id) per vertex, radix-sorted ascending over
//      the id bits and the degree bits (gg_sort.h): descending (degree, id);
//   2. new_id[old] = rank, new offsets = exclusive scan of the sorted
//      degrees;
//   3. every list copied through new_id into its new row (tile expansion,
//      hubs spread over the grid in chunks);
//   4. rows sorted: the rows are in descending degree order, so rows with
//      more than kMediumRow entries are a prefix -- sorted together as
//      (row, value) keys by the radix sort; rows of kSmallRow+1 ..
//      kMediumRow entries by a block bitonic sort in LDS; shorter rows by
//      one thread each (insertion sort in LDS).
#include "gg_expand.h"
#include "gg_sort.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kSmallRow = 32;
constexpr int kMediumRow = 2048;

int BitsFor(uint64_t max_value) {
  int b = 0;
  while (b < 64 && (max_value >> b) != 0) b++;
  return b;
}

__global__ void RlZero(int32_t* p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ __launch_bounds__(kBS) void RlKeys(const int64_t* __restrict__ off, int64_t n, uint64_t* __restrict__ keys) {
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS) {
    const uint64_t deg = static_cast<uint64_t>(off[v + 1] - off[v]);
    keys[v] = ~((deg << 32) | static_cast<uint64_t>(v));
  }
}

// new_id[old] = rank, old_id[rank] = old (old_id may be null);
// row_len[rank] = degree (scanned into offsets next).
__global__ __launch_bounds__(kBS) void RlRanks(const uint64_t* __restrict__ sorted, int64_t n,
                                               int32_t* __restrict__ new_id, int32_t* __restrict__ old_id,
                                               int64_t* __restrict__ row_len) {
  for (int64_t p = blockIdx.x * int64_t(kBS) + threadIdx.x; p < n; p += int64_t(gridDim.x) * kBS) {
    const uint64_t k = ~sorted[p];
    new_id[static_cast<uint32_t>(k)] = static_cast<int32_t>(p);
    if (old_id) old_id[p] = static_cast<int32_t>(static_cast<uint32_t>(k));
    row_len[p] = static_cast<int64_t>(k >> 32);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) row_len[n] = 0;
}

// Entry e of u's list -> the same position of new_id[u]'s list; a weighted
// entry (width 2) keeps its weight.
template <int W>
struct RlCopy {
  const int64_t* old_off;
  const int32_t* old_dst;
  const int64_t* new_off;
  const int32_t* new_id;
  int32_t* new_dst;
  __device__ __forceinline__ void operator()(int32_t u, int32_t, int64_t e, bool valid) {
    if (!valid) return;
    const int64_t to = new_off[new_id[u]] + (e - old_off[u]);
    new_dst[W * to] = new_id[old_dst[W * e]];
    if (W == 2) new_dst[W * to + 1] = old_dst[W * e + 1];
  }
};

template <int W>
__global__ __launch_bounds__(kBS) void RlCopyRows(const int64_t* __restrict__ old_off,
                                                  const int32_t* __restrict__ old_dst,
                                                  const int64_t* __restrict__ new_off,
                                                  const int32_t* __restrict__ new_id, int32_t* new_dst, int64_t n,
                                                  EdgeChunk* chunks, int32_t* nchunks) {
  __shared__ ExpandShared sh;
  RlCopy<W> op{old_off, old_dst, new_off, new_id, new_dst};
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < n; tile += int64_t(gridDim.x) * kBS) {
    const int64_t u = tile + threadIdx.x;
    int64_t b = 0, e = 0;
    if (u < n) {
      b = old_off[u];
      e = old_off[u + 1];
    }
    ExpandTile(sh, static_cast<int32_t>(u), 0, b, e, op, chunks, nchunks, kChunkEdges);
  }
}

template <int W>
__global__ __launch_bounds__(kBS) void RlCopyChunks(const int64_t* __restrict__ old_off,
                                                    const int32_t* __restrict__ old_dst,
                                                    const int64_t* __restrict__ new_off,
                                                    const int32_t* __restrict__ new_id, int32_t* new_dst,
                                                    const EdgeChunk* chunks, const int32_t* nchunks) {
  RlCopy<W> op{old_off, old_dst, new_off, new_id, new_dst};
  ExpandChunks(chunks, *nchunks, op);
}

// bounds[0] = first row with <= kMediumRow entries, bounds[1] = first row
// with <= kSmallRow entries (rows are in descending length order).
__global__ void RlBounds(const int64_t* __restrict__ off, int64_t n, int64_t* bounds) {
  if (threadIdx.x < 2) {
    const int64_t limit = threadIdx.x == 0 ? kMediumRow : kSmallRow;
    int64_t lo = 0, hi = n;  // first p with len(p) <= limit
    while (lo < hi) {
      const int64_t m = lo + ((hi - lo) >> 1);
      if (off[m + 1] - off[m] <= limit)
        hi = m;
      else
        lo = m + 1;
    }
    bounds[threadIdx.x] = lo;
  }
}

// Rows [first, n) with at most kSmallRow entries: one thread per row.
__global__ __launch_bounds__(kBS) void RlSortSmall(const int64_t* __restrict__ off, int32_t* dst,
                                                   const int64_t* __restrict__ bounds, int64_t n) {
  __shared__ int32_t buf[kBS * kSmallRow];
  int32_t* mine = buf + threadIdx.x * kSmallRow;
  for (int64_t r = bounds[1] + blockIdx.x * int64_t(kBS) + threadIdx.x; r < n; r += int64_t(gridDim.x) * kBS) {
    const int64_t b = off[r];
    const int len = static_cast<int>(off[r + 1] - b);
    for (int i = 0; i < len; i++) {
      const int32_t x = dst[b + i];
      int j = i;
      while (j > 0 && mine[j - 1] > x) {
        mine[j] = mine[j - 1];
        j--;
      }
      mine[j] = x;
    }
    for (int i = 0; i < len; i++) dst[b + i] = mine[i];
  }
}

// Rows [bounds[0], bounds[1]): one block per row, bitonic sort in LDS.
__global__ __launch_bounds__(kBS) void RlSortMedium(const int64_t* __restrict__ off, int32_t* dst,
                                                    const int64_t* __restrict__ bounds) {
  __shared__ int32_t a[kMediumRow];
  const int64_t first = bounds[0], last = bounds[1];
  for (int64_t r = first + blockIdx.x; r < last; r += gridDim.x) {
    const int64_t b = off[r];
    const int len = static_cast<int>(off[r + 1] - b);
    int size = 64;
    while (size < len) size <<= 1;
    for (int i = threadIdx.x; i < size; i += kBS) a[i] = i < len ? dst[b + i] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < size; i += kBS) {
          const int p = i ^ j;
          if (p > i) {
            const bool up = (i & k) == 0;
            const int32_t x = a[i], y = a[p];
            if ((x > y) == up) {
              a[i] = y;
              a[p] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < len; i += kBS) dst[b + i] = a[i];
    __syncthreads();
  }
}

// Rows [0, big_rows): keys (row << 32 | value) over the prefix [0, entries).
__global__ __launch_bounds__(kBS) void RlBigKeys(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                 int64_t big_rows, uint64_t* __restrict__ keys) {
  // One block per row, striding over its entries.
  for (int64_t r = blockIdx.x; r < big_rows; r += gridDim.x) {
    const int64_t b = off[r], e = off[r + 1];
    for (int64_t k = b + threadIdx.x; k < e; k += kBS)
      keys[k] = (static_cast<uint64_t>(r) << 32) | static_cast<uint32_t>(dst[k]);
  }
}

__global__ __launch_bounds__(kBS) void RlBigValues(const uint64_t* __restrict__ keys, int64_t m,
                                                   int32_t* __restrict__ dst) {
  for (int64_t k = blockIdx.x * int64_t(kBS) + threadIdx.x; k < m; k += int64_t(gridDim.x) * kBS)
    dst[k] = static_cast<int32_t>(static_cast<uint32_t>(keys[k]));
}

// Steps 1-2 for g's (out-)lists: new_id/old_id (old_id may be null) and
// the offsets of the relabelled lists in new_off (n + 1 entries).
void RankByDegree(gg_graph* g, int32_t* new_id, int32_t* old_id, int64_t* new_off) {
  const int64_t n = g->n;
  const dim3 blk(kBS), grid_n(GridFor(n, kBS, g->cu_count * 16));
  uint64_t* keys = g->Scratch<uint64_t>("rl.keys", n + 1);
  uint64_t* tmp = g->Scratch<uint64_t>("rl.tmp", n + 1);
  const double N = static_cast<double>(n);
  Run(g, "relabel_keys", 16.0 * N, RlKeys, grid_n, blk, static_cast<const int64_t*>(g->out_off), n, keys);
  const int id_bits = BitsFor(n > 0 ? static_cast<uint64_t>(n - 1) : 0);
  const int deg_bits = BitsFor(static_cast<uint64_t>(g->max_degree));
  uint64_t* sorted = RadixSort(g, "rl.sort", keys, tmp, n, 0, id_bits);
  sorted = RadixSort(g, "rl.sort", sorted, sorted == keys ? tmp : keys, n, 32, 32 + deg_bits);
  Run(g, "relabel_ranks", (old_id ? 24.0 : 20.0) * N, RlRanks, grid_n, blk, static_cast<const uint64_t*>(sorted),
      n, new_id, old_id, new_off);
  DeviceScan<int64_t>(g, "rl.offscan", new_off, n + 1, static_cast<int64_t*>(nullptr));
}

// Step 3: every list of g through new_id into new_dst (rows at new_off).
template <int W>
void CopyLists(gg_graph* g, const int64_t* new_off, const int32_t* new_id, int32_t* new_dst) {
  const int64_t n = g->n, entries = g->out_entries;
  const dim3 blk(kBS);
  int32_t* nchunks = g->Scratch<int32_t>("rl.nchunks", 1);
  const bool use_chunks = g->max_degree >= kExpandBS;
  EdgeChunk* chunks = use_chunks ? g->Scratch<EdgeChunk>("rl.chunks", MaxChunks(n, entries)) : nullptr;
  Run(g, "relabel_clear", 0, RlZero, dim3(1), dim3(64), nchunks);
  Run(g, "relabel_copy", 8.0 * n + (8.0 * W + 4.0) * entries, RlCopyRows<W>, dim3(GridFor(n, kBS, g->cu_count * 8)),
      blk, static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst), new_off, new_id,
      new_dst, n, use_chunks ? chunks : static_cast<EdgeChunk*>(nullptr), nchunks);
  if (use_chunks)
    Run(g, "relabel_copy", 0, RlCopyChunks<W>, dim3(g->cu_count * 4), blk, static_cast<const int64_t*>(g->out_off),
        static_cast<const int32_t*>(g->out_dst), new_off, new_id, new_dst, static_cast<const EdgeChunk*>(chunks),
        static_cast<const int32_t*>(nchunks));
}

// Step 4: sorts every row of unweighted lists whose rows are in descending
// length order (rows at off, n rows).
void SortRows(gg_graph* g, const int64_t* off, int32_t* dst) {
  const int64_t n = g->n;
  const dim3 blk(kBS);
  int64_t* bounds = g->Scratch<int64_t>("rl.bounds", 2);
  Run(g, "relabel_bounds", 0, RlBounds, dim3(1), dim3(64), off, n, bounds);
  int64_t hb[2];
  GG_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  int64_t big_entries = 0;
  if (hb[0] > 0) {
    GG_HIP(hipMemcpyAsync(&big_entries, off + hb[0], sizeof(int64_t), hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
  }
  Run(g, "relabel_sort_small", 8.0 * (n - hb[1]), RlSortSmall, dim3(GridFor(n - hb[1], kBS, g->cu_count * 8)), blk,
      off, dst, static_cast<const int64_t*>(bounds), n);
  if (hb[1] > hb[0])
    Run(g, "relabel_sort_medium", 0, RlSortMedium, dim3(GridFor(hb[1] - hb[0], 1, g->cu_count * 8)), blk, off, dst,
        static_cast<const int64_t*>(bounds));
  if (hb[0] > 0) {
    const int id_bits = BitsFor(n > 0 ? static_cast<uint64_t>(n - 1) : 0);
    uint64_t* bk = g->Scratch<uint64_t>("rl.big_keys", big_entries);
    uint64_t* bt = g->Scratch<uint64_t>("rl.big_tmp", big_entries);
    Run(g, "relabel_big_keys", 12.0 * big_entries, RlBigKeys, dim3(GridFor(hb[0], 1, g->cu_count * 8)), blk, off,
        static_cast<const int32_t*>(dst), hb[0], bk);
    uint64_t* s = RadixSort(g, "rl.big", bk, bt, big_entries, 0, id_bits);
    s = RadixSort(g, "rl.big", s, s == bk ? bt : bk, big_entries, 32, 32 + BitsFor(hb[0] - 1));
    Run(g, "relabel_big_values", 12.0 * big_entries, RlBigValues, dim3(GridFor(big_entries, kBS, g->cu_count * 16)),
        blk, static_cast<const uint64_t*>(s), big_entries, dst);
  }
  GG_HIP(hipGetLastError());
}

// out[i] = map[in[i]] over m entries.
__global__ __launch_bounds__(kBS) void RlMapEntries(const int32_t* in, int64_t m, const int32_t* __restrict__ map,
                                                    int32_t* out) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < m; i += int64_t(gridDim.x) * kBS) out[i] = map[in[i]];
}

}  // namespace

const int32_t* UploadOrderLists(gg_graph* g) {
  if (!g->new_id || g->weighted) return g->out_dst;
  if (g->fold_dst) return g->fold_dst;
  const int64_t m = g->out_entries;
  int32_t* f = static_cast<int32_t*>(DeviceAlloc(sizeof(int32_t) * m));
  try {
    // Device ids -> uploaded ids, rows sorted (they are in descending length
    // order, as SortRows needs), uploaded ids -> device ids.
    const dim3 grid(GridFor(m, kBS, g->cu_count * 16)), blk(kBS);
    hipLaunchKernelGGL(RlMapEntries, grid, blk, 0, g->stream, static_cast<const int32_t*>(g->out_dst), m,
                       static_cast<const int32_t*>(g->old_id), f);
    SortRows(g, g->out_off, f);
    hipLaunchKernelGGL(RlMapEntries, grid, blk, 0, g->stream, static_cast<const int32_t*>(f), m,
                       static_cast<const int32_t*>(g->new_id), f);
    GG_HIP(hipGetLastError());
    GG_HIP(hipStreamSynchronize(g->stream));
  } catch (...) {
    (void)hipStreamSynchronize(g->stream);
    DeviceFree(f, sizeof(int32_t) * m);
    throw;
  }
  g->ws.Drop("rl.");
  g->fold_dst = f;
  g->graph_bytes += sizeof(int32_t) * m;
  return f;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_relabel_by_degree(gg_graph* g, gg_graph** out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(out != nullptr, GG_EINVAL, "null output handle");
    *out = nullptr;
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "relabelling needs an unweighted graph upload");
    GG_REQUIRE(!g->directed, GG_ENOTSUP, "Cannot relabel directed graph");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n, entries = g->out_entries;
    GG_REQUIRE(n < (int64_t(1) << 31) && entries < (int64_t(1) << 32), GG_ENOTSUP, "graph too large to relabel");
    gg_graph* r = NewDeviceGraph(g->device, n, entries);
    try {
      int32_t* new_id = g->Scratch<int32_t>("rl.new_id", n + 1);
      BeginTiming(g);
      RankByDegree(g, new_id, nullptr, r->out_off);
      CopyLists<1>(g, r->out_off, new_id, r->out_dst);
      SortRows(g, r->out_off, r->out_dst);
      FinishTiming(g, stats, wt);
      r->max_degree = g->max_degree;
    } catch (...) {
      gg_graph_free(r);
      throw;
    }
    *out = r;
    if (stats) stats->total_ms = wt.ms();
  });
}

extern "C" int gg_graph_order_by_degree(gg_graph* g, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->directed, GG_ENOTSUP, "degree ordering needs an undirected graph");
    GG_REQUIRE(!g->new_id, GG_EINVAL, "graph is already degree-ordered");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n, entries = g->out_entries;
    GG_REQUIRE(entries < (int64_t(1) << 32), GG_ENOTSUP, "graph too large to order");
    const int width = g->weighted ? 2 : 1;
    const size_t off_bytes = sizeof(int64_t) * (n + 1), dst_bytes = sizeof(int32_t) * width * entries;
    const size_t id_bytes = sizeof(int32_t) * (n > 0 ? n : 1);
    int64_t* new_off = nullptr;
    int32_t *new_dst = nullptr, *new_id = nullptr, *old_id = nullptr;
    try {
      new_off = static_cast<int64_t*>(DeviceAlloc(off_bytes));
      new_dst = static_cast<int32_t*>(DeviceAlloc(dst_bytes));
      new_id = static_cast<int32_t*>(DeviceAlloc(id_bytes));
      old_id = static_cast<int32_t*>(DeviceAlloc(id_bytes));
      BeginTiming(g);
      RankByDegree(g, new_id, old_id, new_off);
      if (g->weighted) {
        // Weighted rows keep their uploaded order under the new ids: no
        // kernel result depends on the order within a weighted list.
        CopyLists<2>(g, new_off, new_id, new_dst);
      } else {
        CopyLists<1>(g, new_off, new_id, new_dst);
        SortRows(g, new_off, new_dst);
      }
      FinishTiming(g, stats, wt);
    } catch (...) {
      (void)hipStreamSynchronize(g->stream);
      DeviceFree(new_off, off_bytes);
      DeviceFree(new_dst, dst_bytes);
      DeviceFree(new_id, id_bytes);
      DeviceFree(old_id, id_bytes);
      throw;
    }
    DeviceFree(g->out_off, off_bytes);
    DeviceFree(g->out_dst, dst_bytes);
    g->out_off = g->in_off = new_off;
    g->out_dst = g->in_dst = new_dst;
    g->new_id = new_id;
    g->old_id = old_id;
    g->graph_bytes += 2 * id_bytes;
    // Everything cached from the old layout is rebuilt on demand.
    g->no_in_count = -1;
    g->pr_bins[0] = g->pr_bins[1] = g->pr_bins[2] = -1;
    g->pr_seg[0].built = g->pr_seg[1].built = false;
    g->pr_blocks.built = false;
    g->pk_parts.clear();
    g->ms_heavy = -1;
    g->ms_heavy_chunks = 0;
    g->cc_samples_n = -1;
    g->mb_hubs = -1;
    g->mb_big_hubs = -1;
    if (g->ms_exec) {
      (void)hipGraphExecDestroy(g->ms_exec);
      g->ms_exec = nullptr;
      g->ms_exec_key.clear();
    }
    if (g->sb_exec) {
      (void)hipGraphExecDestroy(g->sb_exec);
      g->sb_exec = nullptr;
      g->sb_exec_key.clear();
    }
    g->ws.Release();
    if (stats) stats->total_ms = wt.ms();
  });
}
