// Delta-stepping single-source shortest paths on gfx950.
//
// Reference: DeltaStep / RelaxEdges (src/sssp.cc:64-153).  Kept: int32
// weights and distances, kDistInf = INT32_MAX / 2 for unreached vertices
// (sssp.cc:64), buckets of width delta processed in increasing order, a
// relaxation as a compare-and-set minimum that re-queues the improved
// vertex in bucket new_dist / delta (sssp.cc:68-85), and processing the
// current bucket again until it stays empty (the effect of the reference's
// bucket fusion, sssp.cc:113-120).  Shortest distances are unique, so the
// result equals the reference's bit for bit whatever the processing order.
//
// GPU design: the current bucket is a device queue F.  A relax launch
// expands F's edges (gg_expand.h, the source distance carried as payload)
// with atomicMin on dist; an improved vertex goes to the `near` queue if it
// stays in the current bucket, else to the `fresh` pile.  When near runs dry
// the bucket is done.  The vertices waiting for later buckets sit in the
// fresh pile, in a `mid` pile holding those inside the current window of
// kWindow buckets, and in a `far` pile holding those beyond it.  The next
// bucket is the one holding the smallest live distance of mid and fresh; a
// split moves its entries into F, keeps the rest of the window in mid and
// appends the rest of fresh to far.  Only when mid and fresh hold nothing
// inside the window is the (large) far pile read: split the same way into F,
// a new window's mid pile and the rest -- once per window, not per bucket.
// Per-vertex stamps keep every queue duplicate-free per launch and every
// pile to one kept and one fresh copy of a vertex, which bounds near and F
// by N and the piles by 2N, so no append can overflow.
#include <climits>

#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kQueueCap = 1024;
constexpr int32_t kDistInf = INT32_MAX / 2;
constexpr int kStripes = 64;
constexpr int64_t kWindow = 8;  // buckets per mid window
// Grid caps of the per-bucket kernels, in blocks per CU.  Bucket frontiers
// and piles are often a few hundred thousand entries; many small blocks keep
// more relaxations and row reads in flight than a few blocks looping over
// tiles.  At Kronecker 26 (4 roots, delta 8) 64 took the step from 703 to
// 294 ms (8 and 4: 703 ms; 16: 501 ms; 32: 328 ms; 128: no better).
constexpr unsigned kBucketBlocksPerCu = 64;
enum { kRelaxed, kImproved, kNumSums };

struct SsspCounters {
  int32_t near_tail;   // relax output (same bucket)
  int32_t fresh_tail;  // relax output (later buckets)
  int32_t f_tail;      // split output (next bucket)
  int32_t mid2_tail;   // split output (the window's entries)
  int32_t fout_tail;   // split output (entries beyond the window)
  int32_t min_far;     // smallest pile distance >= bucket end
  int32_t nchunks;
  unsigned long long sums[kStripes][kNumSums];
};

__global__ __launch_bounds__(kBS) void SsspReset(SsspCounters* c) {
  unsigned long long* p = &c->sums[0][0];
  for (int i = threadIdx.x; i < kStripes * kNumSums; i += blockDim.x) p[i] = 0;
  if (threadIdx.x == 0) {
    c->near_tail = c->fresh_tail = c->f_tail = c->mid2_tail = c->fout_tail = 0;
    c->min_far = kDistInf;
    c->nchunks = 0;
  }
}

__global__ __launch_bounds__(kBS) void SsspInit(int32_t* __restrict__ dist, int32_t* __restrict__ stamp,
                                                int32_t* __restrict__ freshstamp, int32_t* __restrict__ midstamp,
                                                int32_t* __restrict__ farstamp, int64_t n, int32_t src,
                                                int32_t* __restrict__ frontier) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    dist[i] = i == src ? 0 : kDistInf;
    stamp[i] = -1;
    freshstamp[i] = -1;
    midstamp[i] = -1;
    farstamp[i] = -1;
    if (i == src) frontier[0] = src;
  }
}

enum PublishMode { kAfterRelax = 0, kAfterMin = 1, kAfterSplit = 2 };
enum { kMailNear, kMailFresh, kMailF, kMailMid, kMailFout, kMailMin, kMailRelaxed, kMailImproved, kMailValues };

// Posts the counters the host needs after each kind of step and resets the
// ones the next step appends to.  One block.
__global__ __launch_bounds__(kBS) void SsspPublish(SsspCounters* c, int mode, unsigned long long* mb,
                                                   unsigned long long seq) {
  __shared__ unsigned long long tot[kNumSums];
  if (threadIdx.x < kNumSums) tot[threadIdx.x] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < kStripes * kNumSums; i += kBS) {
    unsigned long long* p = &c->sums[i / kNumSums][i % kNumSums];
    const unsigned long long v = *p;
    if (v) atomicAdd(&tot[i % kNumSums], v);
    *p = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long vals[kMailValues];
    vals[kMailNear] = static_cast<unsigned long long>(c->near_tail);
    vals[kMailFresh] = static_cast<unsigned long long>(c->fresh_tail);
    vals[kMailF] = static_cast<unsigned long long>(c->f_tail);
    vals[kMailMid] = static_cast<unsigned long long>(c->mid2_tail);
    vals[kMailFout] = static_cast<unsigned long long>(c->fout_tail);
    vals[kMailMin] = static_cast<unsigned long long>(c->min_far);
    vals[kMailRelaxed] = tot[kRelaxed];
    vals[kMailImproved] = tot[kImproved];
    if (mode == kAfterRelax) {
      c->near_tail = 0;
      c->nchunks = 0;
    } else if (mode == kAfterMin) {
      c->min_far = kDistInf;
    } else {  // a split consumed the fresh pile; the host takes its outputs' sizes
      c->fresh_tail = c->f_tail = c->mid2_tail = c->fout_tail = 0;
    }
    MailPost(mb, vals, kMailValues, seq);
  }
}

// Relaxes edge e of u (payload = dist[u] when u was loaded).
struct RelaxVisit {
  const int2* wdst;
  int32_t* dist;
  int32_t* stamp;
  int32_t* freshstamp;
  int32_t bucket_end;
  int32_t near_tag, fresh_tag;
  LdsQueue<kQueueCap>* lnear;
  LdsQueue<kQueueCap>* lfresh;
  int32_t* near;
  int32_t* fresh;
  SsspCounters* c;
  unsigned long long relaxed = 0, improved = 0;

  __device__ __forceinline__ void operator()(int32_t, int32_t du, int64_t e, bool valid) {
    bool to_near = false, to_fresh = false;
    int32_t v = 0;
    if (valid) {
      const int2 vw = wdst[e];
      v = vw.x;
      const int64_t nd = static_cast<int64_t>(du) + vw.y;
      relaxed++;
      if (nd < kDistInf && nd < dist[v]) {
        const int32_t d = static_cast<int32_t>(nd);
        const int32_t old = atomicMin(&dist[v], d);
        if (d < old) {
          improved++;
          if (d < bucket_end)
            to_near = atomicExch(&stamp[v], near_tag) != near_tag;
          else
            to_fresh = atomicExch(&freshstamp[v], fresh_tag) != fresh_tag;
        }
      }
    }
    LdsQueuePush(*lnear, to_near, v, near, &c->near_tail);
    LdsQueuePush(*lfresh, to_fresh, v, fresh, &c->fresh_tail);
  }
};

__device__ __forceinline__ void FlushRelax(RelaxVisit& rv, SsspCounters* c, unsigned long long* s_red) {
  LdsQueueFlush(*rv.lnear, rv.near, &c->near_tail);
  LdsQueueFlush(*rv.lfresh, rv.fresh, &c->fresh_tail);
  const unsigned long long r = BlockSum<kBS>(rv.relaxed, s_red);
  const unsigned long long im = BlockSum<kBS>(rv.improved, s_red);
  if (threadIdx.x == 0) {
    unsigned long long* st = c->sums[blockIdx.x & (kStripes - 1)];
    if (r) atomicAdd(&st[kRelaxed], r);
    if (im) atomicAdd(&st[kImproved], im);
  }
}

__global__ __launch_bounds__(kBS) void SsspRelax(const int64_t* __restrict__ off, const int2* __restrict__ wdst,
                                                 int32_t* __restrict__ dist, int32_t* __restrict__ stamp,
                                                 int32_t* __restrict__ freshstamp,
                                                 const int32_t* __restrict__ frontier, int32_t fn,
                                                 int32_t bucket_end, int32_t near_tag, int32_t fresh_tag,
                                                 int32_t* __restrict__ near, int32_t* __restrict__ fresh,
                                                 SsspCounters* c, EdgeChunk* chunks, int64_t chunk_min) {
  __shared__ ExpandShared sh;
  __shared__ LdsQueue<kQueueCap> lnear, lfresh;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lnear);
  LdsQueueInit(lfresh);
  RelaxVisit rv{wdst, dist, stamp, freshstamp, bucket_end, near_tag, fresh_tag, &lnear, &lfresh, near, fresh, c};
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < fn; tile += int64_t(gridDim.x) * kBS) {
    const int64_t i = tile + threadIdx.x;
    int32_t u = 0, du = 0;
    int64_t beg = 0, end = 0;
    if (i < fn) {
      u = frontier[i];
      du = dist[u];
      beg = off[u];
      end = off[u + 1];
    }
    ExpandTile(sh, u, du, beg, end, rv, chunks, &c->nchunks, chunk_min);
  }
  FlushRelax(rv, c, s_red);
}

__global__ __launch_bounds__(kBS) void SsspRelaxChunks(const int2* __restrict__ wdst, int32_t* __restrict__ dist,
                                                       int32_t* __restrict__ stamp,
                                                       int32_t* __restrict__ freshstamp, int32_t bucket_end,
                                                       int32_t near_tag, int32_t fresh_tag,
                                                       int32_t* __restrict__ near, int32_t* __restrict__ fresh,
                                                       SsspCounters* c, const EdgeChunk* chunks) {
  __shared__ LdsQueue<kQueueCap> lnear, lfresh;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lnear);
  LdsQueueInit(lfresh);
  RelaxVisit rv{wdst, dist, stamp, freshstamp, bucket_end, near_tag, fresh_tag, &lnear, &lfresh, near, fresh, c};
  ExpandChunks(chunks, c->nchunks, rv);
  FlushRelax(rv, c, s_red);
}

// Smallest distance >= bucket_end among the entries of two piles.
__global__ __launch_bounds__(kBS) void SsspPileMin(const int32_t* __restrict__ pa, int32_t na,
                                                   const int32_t* __restrict__ pb, int32_t nb,
                                                   const int32_t* __restrict__ dist, int32_t bucket_end,
                                                   SsspCounters* c) {
  __shared__ int32_t s_min[kBS / kWave];
  int32_t m = kDistInf;
  const int64_t total = int64_t(na) + nb;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < total; i += int64_t(gridDim.x) * kBS) {
    const int32_t d = dist[i < na ? pa[i] : pb[i - na]];
    if (d >= bucket_end && d < m) m = d;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(m, o);
    m = t < m ? t : m;
  }
  if (LaneId() == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBS / kWave; w++) m = s_min[w] < m ? s_min[w] : m;
    m = s_min[0] < m ? s_min[0] : m;
    if (m < kDistInf) atomicMin(&c->min_far, m);
  }
}

// Splits the entries of two piles: distance in [old_end, new_end) -> the
// next frontier, [new_end, win_end) -> mid2, the rest -> fout (each once,
// by its stamp); entries below old_end were settled in an earlier bucket
// and are dropped.
__global__ __launch_bounds__(kBS) void SsspSplit(const int32_t* __restrict__ pa, int32_t na,
                                                 const int32_t* __restrict__ pb, int32_t nb,
                                                 const int32_t* __restrict__ dist, int32_t* __restrict__ stamp,
                                                 int32_t* __restrict__ midstamp, int32_t* __restrict__ farstamp,
                                                 int32_t old_end, int32_t new_end, int32_t win_end, int32_t f_tag,
                                                 int32_t mid_tag, int32_t far_tag, int32_t* __restrict__ frontier,
                                                 int32_t* __restrict__ mid2, int32_t* __restrict__ fout,
                                                 SsspCounters* c) {
  __shared__ LdsQueue<kQueueCap> lf, lm, lk;
  LdsQueueInit(lf);
  LdsQueueInit(lm);
  LdsQueueInit(lk);
  const int64_t total = int64_t(na) + nb;
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < total; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    bool to_f = false, to_m = false, keep = false;
    int32_t v = 0;
    if (i < total) {
      v = i < na ? pa[i] : pb[i - na];
      const int32_t d = dist[v];
      if (d >= old_end) {
        if (d < new_end)
          to_f = atomicExch(&stamp[v], f_tag) != f_tag;
        else if (d < win_end)
          to_m = atomicExch(&midstamp[v], mid_tag) != mid_tag;
        else
          keep = atomicExch(&farstamp[v], far_tag) != far_tag;
      }
    }
    LdsQueuePush(lf, to_f, v, frontier, &c->f_tail);
    LdsQueuePush(lm, to_m, v, mid2, &c->mid2_tail);
    LdsQueuePush(lk, keep, v, fout, &c->fout_tail);
  }
  LdsQueueFlush(lf, frontier, &c->f_tail);
  LdsQueueFlush(lm, mid2, &c->mid2_tail);
  LdsQueueFlush(lk, fout, &c->fout_tail);
}

// Reached vertices and their summed out-degree (for GTEPS).
__global__ __launch_bounds__(kBS) void SsspReach(const int32_t* __restrict__ dist, const int64_t* __restrict__ off,
                                                 int64_t n, unsigned long long* out) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long cnt = 0, deg = 0;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    if (dist[i] < kDistInf) {
      cnt++;
      deg += static_cast<unsigned long long>(off[i + 1] - off[i]);
    }
  cnt = BlockSum<kBS>(cnt, s_red);
  deg = BlockSum<kBS>(deg, s_red);
  if (threadIdx.x == 0) {
    if (cnt) atomicAdd(&out[0], cnt);
    if (deg) atomicAdd(&out[1], deg);
  }
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_sssp(gg_graph* g, int32_t source, int32_t delta, int32_t* dist_out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(g->weighted, GG_ENOTSUP, "SSSP needs a weighted graph upload");
    GG_REQUIRE(source >= 0 && source < g->n, GG_EINVAL, "source out of range");
    GG_REQUIRE(delta > 0, GG_EINVAL, "delta must be positive");
    GG_HIP(hipSetDevice(g->device));
    source = DeviceVertex(g, source);
    const int64_t n = g->n;
    hipStream_t st = g->stream;
    const int2* wdst = reinterpret_cast<const int2*>(g->out_dst);

    int32_t* dist = g->Scratch<int32_t>("sssp.dist", n);
    int32_t* stamp = g->Scratch<int32_t>("sssp.stamp", n);
    int32_t* freshstamp = g->Scratch<int32_t>("sssp.freshstamp", n);
    int32_t* midstamp = g->Scratch<int32_t>("sssp.midstamp", n);
    int32_t* farstamp = g->Scratch<int32_t>("sssp.farstamp", n);
    int32_t* fq = g->Scratch<int32_t>("sssp.f", n + 64);
    int32_t* nq = g->Scratch<int32_t>("sssp.near", n + 64);
    int32_t* fresh = g->Scratch<int32_t>("sssp.fresh", n + 64);
    int32_t* mid = g->Scratch<int32_t>("sssp.mid", 2 * n + 64);
    int32_t* mid2 = g->Scratch<int32_t>("sssp.mid2", 2 * n + 64);
    int32_t* far = g->Scratch<int32_t>("sssp.far", 2 * n + 64);
    int32_t* far2 = g->Scratch<int32_t>("sssp.far2", 2 * n + 64);
    SsspCounters* dc = g->Scratch<SsspCounters>("sssp.ctr", 1);
    const bool use_chunks = g->max_degree >= kExpandBS;
    EdgeChunk* chunks =
        use_chunks ? g->Scratch<EdgeChunk>("sssp.chunks", MaxChunks(n, g->out_entries)) : nullptr;

    const dim3 blk(kBS);
    const dim3 grid_n(GridFor(n, kBS, g->cu_count * 32));
    const double N = static_cast<double>(n);
    int64_t launches = 0;
    unsigned long long mail[kMailValues];
    auto publish = [&](int mode) {
      const unsigned long long seq = MailNext(g);
      Run(g, "sssp_publish", 0, SsspPublish, dim3(1), blk, dc, mode, g->mail_dev, seq);
      launches++;
      MailWait(g, seq, mail, kMailValues);
    };

    BeginTiming(g);
    Run(g, "sssp_reset", 0, SsspReset, dim3(1), blk, dc);
    Run(g, "sssp_init", 20 * N, SsspInit, grid_n, blk, dist, stamp, freshstamp, midstamp, farstamp, n, source, fq);
    launches += 2;
    auto clamp = [](int64_t x) { return static_cast<int32_t>(x < kDistInf ? x : kDistInf); };
    int64_t fsize = 1, fresh_size = 0, mid_size = 0, far_size = 0;
    int64_t bucket_end = delta;
    int64_t window_end = kWindow * static_cast<int64_t>(delta);
    int32_t tag = 0;
    // A far stamp holds until the far pile is split: one appended copy of a
    // vertex per window besides the copy kept by the last split.
    int32_t far_tag = ++tag;
    int64_t relax_launches = 0, buckets = 1, relaxed = 0;
    // Each relax launch strictly lowers some distance or empties the queue,
    // so this bound is never reached by a correct run; it only turns a
    // logic error into an error instead of an endless loop.
    const int64_t max_launches = 64 * n + 100000;
    // Smallest live distance (>= bucket_end) over two piles, kDistInf if none.
    auto pile_min = [&](const int32_t* pa, int64_t na, const int32_t* pb, int64_t nb) {
      const int64_t total = na + nb;
      if (total == 0) return static_cast<int64_t>(kDistInf);
      Run(g, "sssp_pile_min", 8.0 * total, SsspPileMin, dim3(GridFor(total, kBS, g->cu_count * kBucketBlocksPerCu)), blk, pa,
          static_cast<int32_t>(na), pb, static_cast<int32_t>(nb), static_cast<const int32_t*>(dist),
          clamp(bucket_end), dc);
      launches++;
      publish(kAfterMin);
      return static_cast<int64_t>(mail[kMailMin]);
    };
    while (true) {
      const int32_t fresh_tag = ++tag;
      while (fsize > 0) {
        GG_REQUIRE(relax_launches < max_launches, GG_EHIP, "SSSP did not converge");
        const int32_t near_tag = ++tag;
        const int32_t end32 = clamp(bucket_end);
        const int64_t chunk_min = fsize <= 4096 ? kExpandBS : kChunkEdges;
        const int idx = Run(g, "sssp_relax", 12.0 * fsize, SsspRelax,
                            dim3(GridFor(fsize, kBS, g->cu_count * kBucketBlocksPerCu)), blk, g->out_off, wdst, dist,
                            stamp, freshstamp, static_cast<const int32_t*>(fq),
                            static_cast<int32_t>(fsize), end32, near_tag, fresh_tag, nq, fresh, dc,
                            use_chunks ? chunks : static_cast<EdgeChunk*>(nullptr), chunk_min);
        launches++;
        if (use_chunks) {
          Run(g, "sssp_relax_chunks", 0, SsspRelaxChunks, dim3(g->cu_count * kBucketBlocksPerCu), blk, wdst, dist,
              stamp, freshstamp, end32, near_tag, fresh_tag, nq, fresh, dc,
              static_cast<const EdgeChunk*>(chunks));
          launches++;
        }
        publish(kAfterRelax);
        relax_launches++;
        relaxed += static_cast<int64_t>(mail[kMailRelaxed]);
        CreditBytes(g, idx, 8.0 * mail[kMailRelaxed] + 8.0 * mail[kMailImproved]);
        fsize = static_cast<int64_t>(mail[kMailNear]);
        fresh_size = static_cast<int64_t>(mail[kMailFresh]);
        std::swap(fq, nq);
      }
      // The bucket is done.  Next: the smallest live distance of mid and
      // fresh when it lies inside the window (far holds only distances at or
      // beyond the window's end), else a new window from far and fresh.
      GG_REQUIRE(buckets < max_launches, GG_EHIP, "SSSP did not converge");
      int64_t m = pile_min(mid, mid_size, fresh, fresh_size);
      // (kDistInf, "no live entry", can lie below a window end past it.)
      const bool in_window = m < kDistInf && m < window_end;
      if (!in_window) {
        m = pile_min(far, far_size, fresh, fresh_size);
        if (m >= kDistInf) break;
      }
      const int64_t next_end = (m / delta + 1) * static_cast<int64_t>(delta);
      if (!in_window) window_end = next_end + (kWindow - 1) * static_cast<int64_t>(delta);
      const int32_t f_tag = ++tag, mid_tag = ++tag;
      // A window split appends beyond-window entries to far; a new window
      // rewrites far (the old mid holds no live entry then).
      const int32_t keep_far = in_window ? far_tag : ++tag;
      const int32_t* pa = in_window ? mid : far;
      const int64_t na = in_window ? mid_size : far_size;
      int32_t* fout = in_window ? far + far_size : far2;
      const int64_t total = na + fresh_size;
      Run(g, in_window ? "sssp_mid_split" : "sssp_far_split", 12.0 * total, SsspSplit,
          dim3(GridFor(total, kBS, g->cu_count * kBucketBlocksPerCu)), blk, pa, static_cast<int32_t>(na),
          static_cast<const int32_t*>(fresh), static_cast<int32_t>(fresh_size), static_cast<const int32_t*>(dist),
          stamp, midstamp, farstamp, clamp(bucket_end), clamp(next_end), clamp(window_end), f_tag, mid_tag, keep_far,
          fq, mid2, fout, dc);
      launches++;
      publish(kAfterSplit);
      fsize = static_cast<int64_t>(mail[kMailF]);
      mid_size = static_cast<int64_t>(mail[kMailMid]);
      fresh_size = 0;
      std::swap(mid, mid2);
      if (in_window) {
        far_size += static_cast<int64_t>(mail[kMailFout]);
      } else {
        far_size = static_cast<int64_t>(mail[kMailFout]);
        std::swap(far, far2);
        far_tag = ++tag;
      }
      GG_REQUIRE(far_size <= 2 * n + 64 && mid_size <= 2 * n + 64, GG_EHIP, "SSSP pile bound exceeded");
      bucket_end = next_end;
      buckets++;
    }
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (dist_out) CopyOutVertexI32(g, dist, dist_out, false);
    if (stats) {
      unsigned long long* acc = g->Scratch<unsigned long long>("sssp.reach", 2);
      GG_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(unsigned long long), st));
      SsspReach<<<grid_n, blk, 0, st>>>(dist, g->out_off, n, acc);
      unsigned long long h[2];
      GG_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      stats->total_ms = wt.ms();
      stats->iterations = buckets;
      stats->td_steps = relax_launches;
      stats->reached = static_cast<int64_t>(h[0]);
      stats->reached_edges = static_cast<int64_t>(h[1]);
      stats->edges_examined = relaxed;
      stats->launches = launches;
    }
  });
}
