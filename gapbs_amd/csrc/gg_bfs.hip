// Direction-optimizing BFS on gfx950.
//
// Reference: DOBFS and its steps (src/bfs.cc:46-179).  What is kept:
//   * parent[] encodes unvisited vertices as -out_degree (or -1 for degree
//     0), so a top-down step sums the degrees of the vertices it discovers
//     without reading offsets (bfs.cc:31-35, 114-120, 80-82);
//   * the controller: go bottom-up when scout > edges_to_check / alpha, stay
//     bottom-up while awake >= old_awake || awake > N / beta, subtract scout
//     from edges_to_check on top-down steps only, reset scout to 1 after a
//     bottom-up phase (bfs.cc:142-173) -- so the GPU takes exactly the
//     reference's sequence of directions;
//   * bottom-up scans each unvisited vertex's in-neighbours in sorted order
//     and takes the first one in the frontier (bfs.cc:50-62), so bottom-up
//     parents are the reference's;
//   * the final fix-up parent < -1 -> -1 (bfs.cc:174-177).
// What is GPU-specific:
//   * top-down expands 256-vertex tiles of the frontier with gg_expand.h
//     (hubs cut into chunks spread over the grid; when the frontier is small
//     every vertex with >= 256 edges is chunked); discovered vertices are
//     staged in an LDS queue and flushed with one atomic per block;
//   * bottom-up gives each wavefront 64 consecutive vertices, so the next
//     frontier word is one ballot, written without atomics; a visited bitmap
//     (with vertices that have no in-neighbours pre-set) lets whole words be
//     skipped; neighbours are probed 2, then 8 at a time;
//   * per-step sums are striped over 64 slots and posted to the host by a
//     one-block kernel through a mapped mailbox; the host runs the
//     reference's controller on them.
#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kQueueCap = 2048;  // LDS-staged queue entries per block

// Sums are striped over kStripes slots (slot = block index mod kStripes)
// so that thousands of blocks do not serialise on one address.
constexpr int kStripes = 64;
enum { kScout, kAwake, kReach, kExamined, kVexam, kNumSums };

struct BfsCounters {
  int32_t qtail;      // top-down output queue tail
  int32_t nchunks;    // top-down hub chunks emitted this step
  int32_t src_no_in;  // set by BfsInit: the source has no in-neighbours
  int32_t bmtail;     // bitmap-to-queue output tail
  unsigned long long sums[kStripes][kNumSums];
};

__device__ __forceinline__ unsigned long long* Stripe(BfsCounters* c) {
  return c->sums[blockIdx.x & (kStripes - 1)];
}

__global__ __launch_bounds__(kBS) void BfsResetCounters(BfsCounters* c) {
  unsigned long long* p = &c->sums[0][0];
  for (int i = threadIdx.x; i < kStripes * kNumSums; i += blockDim.x) p[i] = 0;
  if (threadIdx.x == 0) {
    c->qtail = 0;
    c->nchunks = 0;
    c->src_no_in = 0;
    c->bmtail = 0;
  }
}

// Values BfsPublish posts to the host mailbox.
enum { kMailQtail = kNumSums, kMailSrcNoIn, kMailValues };

// Sums the striped counters, posts the totals to the host mailbox and
// clears the per-step counters.  One block of kBS threads.
__global__ __launch_bounds__(kBS) void BfsPublish(BfsCounters* c, unsigned long long* mb,
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
    for (int k = 0; k < kNumSums; k++) vals[k] = tot[k];
    vals[kMailQtail] = static_cast<unsigned long long>(c->qtail);
    vals[kMailSrcNoIn] = static_cast<unsigned long long>(c->src_no_in);
    c->qtail = 0;
    c->nchunks = 0;
    c->bmtail = 0;
    MailPost(mb, vals, kMailValues, seq);
  }
}

// Bit l of word w: vertex 64w + l has no in-neighbours (or lies past n).
__global__ __launch_bounds__(kBS) void BfsNoInNeighbours(const int64_t* __restrict__ in_off, int64_t n,
                                                         uint64_t* __restrict__ no_in,
                                                         unsigned long long* count) {
  __shared__ unsigned long long s_red[kBS / kWave];
  const int lane = LaneId();
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  unsigned long long cnt = 0;
  for (int64_t w = wave; w < nwords; w += nwaves) {
    const int64_t u = w * 64 + lane;
    const uint64_t m = __ballot(u >= n || in_off[u + 1] == in_off[u]);
    if (lane == 0) {
      no_in[w] = m;
      const uint64_t valid = (w == nwords - 1 && (n & 63)) ? ((1ull << (n & 63)) - 1) : ~0ull;
      cnt += __popcll(m & valid);
    }
  }
  cnt = BlockSum<kBS>(cnt, s_red);
  if (threadIdx.x == 0 && cnt) atomicAdd(count, cnt);
}

__global__ __launch_bounds__(kBS) void BfsInit(const int64_t* __restrict__ off,
                                               const int64_t* __restrict__ in_off,
                                               int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                               int64_t n, int32_t src, int32_t* __restrict__ queue,
                                               BfsCounters* c) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    const int64_t deg = off[i + 1] - off[i];
    int32_t p = deg != 0 ? static_cast<int32_t>(-deg) : -1;
    if (i == src) {
      p = src;
      queue[0] = src;
      c->sums[0][kScout] = static_cast<unsigned long long>(deg);
      c->sums[0][kReach] = static_cast<unsigned long long>(deg);
      c->src_no_in = in_off[i + 1] == in_off[i];
    }
    parent[i] = p;
    if (depth) depth[i] = (i == src) ? 0 : -1;
  }
}

// One frontier edge u -> dst[e] (valid lanes only act; all lanes of a
// wavefront call together).  Winners of the parent CAS join the next queue.
struct TdVisit {
  const int32_t* dst;
  int32_t* parent;
  int32_t* depth;
  const int64_t* off;
  LdsQueue<kQueueCap>* lq;
  int32_t* qout;
  int32_t* qtail;
  int32_t level;
  unsigned long long scout = 0, reach = 0;

  __device__ __forceinline__ void operator()(int32_t u, int32_t, int64_t e, bool valid) {
    bool won = false;
    int32_t cur = 0, v = 0;
    if (valid) {
      v = dst[e];
      cur = parent[v];
      if (cur < 0) won = atomicCAS(&parent[v], cur, u) == cur;
    }
    if (won) {
      if (depth) depth[v] = level;
      scout += static_cast<unsigned long long>(-static_cast<int64_t>(cur));
      // -1 encodes both degree 0 and degree 1; read the degree for those.
      reach += cur < -1 ? static_cast<unsigned long long>(-static_cast<int64_t>(cur))
                        : static_cast<unsigned long long>(off[v + 1] - off[v]);
    }
    LdsQueuePush(*lq, won, v, qout, qtail);
  }
};

__device__ __forceinline__ void FlushTdCounters(const TdVisit& visit, unsigned long long examined,
                                                BfsCounters* c, unsigned long long* s_red) {
  const unsigned long long sc = BlockSum<kBS>(visit.scout, s_red);
  const unsigned long long re = BlockSum<kBS>(visit.reach, s_red);
  const unsigned long long ex = BlockSum<kBS>(examined, s_red);
  if (threadIdx.x == 0) {
    unsigned long long* st = Stripe(c);
    if (sc) atomicAdd(&st[kScout], sc);
    if (re) atomicAdd(&st[kReach], re);
    if (ex) atomicAdd(&st[kExamined], ex);
  }
}

__global__ __launch_bounds__(kBS) void BfsTopDown(const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ dst,
                                                  int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                                  const int32_t* __restrict__ qin, int32_t qn,
                                                  int32_t* __restrict__ qout, BfsCounters* c, int32_t level,
                                                  EdgeChunk* chunks, int64_t chunk_min) {
  __shared__ ExpandShared sh;
  __shared__ LdsQueue<kQueueCap> lq;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lq);
  TdVisit visit{dst, parent, depth, off, &lq, qout, &c->qtail, level};
  unsigned long long examined = 0;
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < qn; tile += int64_t(gridDim.x) * kBS) {
    const int64_t i = tile + threadIdx.x;
    int32_t u = 0;
    int64_t beg = 0, end = 0;
    if (i < qn) {
      u = qin[i];
      beg = off[u];
      end = off[u + 1];
      examined += static_cast<unsigned long long>(end - beg);
    }
    ExpandTile(sh, u, 0, beg, end, visit, chunks, &c->nchunks, chunk_min);
  }
  LdsQueueFlush(lq, qout, &c->qtail);
  FlushTdCounters(visit, examined, c, s_red);
}

// Edges of the vertices BfsTopDown cut into chunks.
__global__ __launch_bounds__(kBS) void BfsTopDownChunks(const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ dst,
                                                        int32_t* __restrict__ parent,
                                                        int32_t* __restrict__ depth,
                                                        int32_t* __restrict__ qout, BfsCounters* c,
                                                        int32_t level, const EdgeChunk* chunks) {
  __shared__ LdsQueue<kQueueCap> lq;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lq);
  TdVisit visit{dst, parent, depth, off, &lq, qout, &c->qtail, level};
  ExpandChunks(chunks, c->nchunks, visit);
  LdsQueueFlush(lq, qout, &c->qtail);
  FlushTdCounters(visit, 0, c, s_red);
}

__global__ __launch_bounds__(kBS) void FillU64(uint64_t* __restrict__ p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) p[i] = v;
}

__global__ __launch_bounds__(kBS) void BfsQueueToBitmap(const int32_t* __restrict__ q, int32_t qn,
                                                        unsigned long long* __restrict__ bm) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < qn; i += int64_t(gridDim.x) * kBS) {
    const int32_t v = q[i];
    atomicOr(&bm[v >> 6], 1ull << (v & 63));
  }
}

// visited[w] bit l <=> parent[64w + l] >= 0, or vertex 64w + l has no
// in-neighbours (bottom-up can never discover it), or lies past n.  Rebuilt
// at each switch to bottom-up; bottom-up steps keep it current.
__global__ __launch_bounds__(kBS) void BfsVisitedFromParent(const int32_t* __restrict__ parent, int64_t n,
                                                            const uint64_t* __restrict__ no_in,
                                                            uint64_t* __restrict__ visited) {
  const int lane = LaneId();
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  for (int64_t w = wave; w < nwords; w += nwaves) {
    const int64_t u = w * 64 + lane;
    const uint64_t m = __ballot(u >= n || parent[u] >= 0);
    if (lane == 0) visited[w] = m | no_in[w];
  }
}

// Bottom-up step.  Each wavefront owns bitmap words; lane l handles vertex
// 64*word + l and scans its in-neighbours in sorted order, taking the first
// one in `front`.
constexpr int kBuUnroll = 8;

__global__ __launch_bounds__(kBS) void BfsBottomUp(const int64_t* __restrict__ in_off,
                                                   const int32_t* __restrict__ in_dst,
                                                   const int64_t* __restrict__ out_off,
                                                   int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                                   const uint64_t* __restrict__ front,
                                                   uint64_t* __restrict__ next, uint64_t* __restrict__ visited,
                                                   int64_t n, BfsCounters* c, int32_t level) {
  __shared__ unsigned long long s_red[kBS / kWave];
  const int lane = LaneId();
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  unsigned long long awake = 0, reach = 0, examined = 0, vexam = 0;
  for (int64_t w = wave; w < nwords; w += nwaves) {
    const uint64_t vis = visited[w];
    if (vis == ~0ull) {
      if (lane == 0) next[w] = 0;
      continue;
    }
    const int64_t u = w * 64 + lane;
    bool found = false;
    if (!((vis >> lane) & 1)) {
      vexam++;
      const int64_t b = in_off[u], e = in_off[u + 1];
      int64_t hit = -1;
      int32_t hv = 0;
      // Most vertices that will be found are found at their first or second
      // in-neighbour: probe two, then eight at a time (independent loads).
      {
        const int32_t v0 = b < e ? in_dst[b] : -1;
        const int32_t v1 = b + 1 < e ? in_dst[b + 1] : -1;
        const uint64_t w0 = v0 >= 0 ? front[v0 >> 6] : 0;
        const uint64_t w1 = v1 >= 0 ? front[v1 >> 6] : 0;
        if (v1 >= 0 && ((w1 >> (v1 & 63)) & 1)) {
          hit = b + 1;
          hv = v1;
        }
        if (v0 >= 0 && ((w0 >> (v0 & 63)) & 1)) {
          hit = b;
          hv = v0;
        }
      }
      for (int64_t k = b + 2; k < e && hit < 0; k += kBuUnroll) {
        int32_t v[kBuUnroll];
        uint64_t word[kBuUnroll];
#pragma unroll
        for (int j = 0; j < kBuUnroll; j++) v[j] = (k + j < e) ? in_dst[k + j] : -1;
#pragma unroll
        for (int j = 0; j < kBuUnroll; j++) word[j] = v[j] >= 0 ? front[v[j] >> 6] : 0;
#pragma unroll
        for (int j = kBuUnroll - 1; j >= 0; j--)
          if (v[j] >= 0 && ((word[j] >> (v[j] & 63)) & 1)) {
            hit = k + j;
            hv = v[j];
          }
      }
      if (hit >= 0) {
        found = true;
        parent[u] = hv;
        if (depth) depth[u] = level;
        reach += static_cast<unsigned long long>(out_off[u + 1] - out_off[u]);
        examined += static_cast<unsigned long long>(hit + 1 - b);
      } else {
        examined += static_cast<unsigned long long>(e - b);
      }
    }
    const uint64_t m = __ballot(found);
    if (lane == 0) {
      next[w] = m;
      visited[w] = vis | m;
      awake += __popcll(m);
    }
  }
  const unsigned long long aw = BlockSum<kBS>(awake, s_red);
  const unsigned long long re = BlockSum<kBS>(reach, s_red);
  const unsigned long long ex = BlockSum<kBS>(examined, s_red);
  const unsigned long long vx = BlockSum<kBS>(vexam, s_red);
  if (threadIdx.x == 0) {
    unsigned long long* st = Stripe(c);
    if (aw) atomicAdd(&st[kAwake], aw);
    if (re) atomicAdd(&st[kReach], re);
    if (ex) atomicAdd(&st[kExamined], ex);
    if (vx) atomicAdd(&st[kVexam], vx);
  }
}

// Compacts the set bits of a bitmap into a queue.  Each block owns a
// contiguous range of words: it counts its bits, reserves its output range
// with one atomic, then writes the vertices.
__global__ __launch_bounds__(kBS) void BfsBitmapToQueue(const uint64_t* __restrict__ bm, int64_t nwords,
                                                        int32_t* __restrict__ q, int32_t* tail) {
  __shared__ int s_scan[kBS / kWave];
  __shared__ int s_red[kBS / kWave];
  __shared__ int s_base;
  const int64_t per = (nwords + gridDim.x - 1) / gridDim.x;
  const int64_t w0 = int64_t(blockIdx.x) * per;
  const int64_t w1 = w0 + per < nwords ? w0 + per : nwords;
  int cnt = 0;
  for (int64_t w = w0 + threadIdx.x; w < w1; w += kBS) cnt += __popcll(bm[w]);
  cnt = BlockSum<kBS>(cnt, s_red);
  if (cnt == 0) return;  // uniform: every thread holds the block total
  if (threadIdx.x == 0) s_base = atomicAdd(tail, cnt);
  __syncthreads();
  int run = s_base;
  for (int64_t wb = w0; wb < w1; wb += kBS) {
    const int64_t w = wb + threadIdx.x;
    uint64_t word = w < w1 ? bm[w] : 0;
    int total = 0;
    int at = run + BlockExclusiveScan<kBS>(__popcll(word), &total, s_scan);
    while (word) {
      q[at++] = static_cast<int32_t>(w * 64 + __builtin_ctzll(word));
      word &= word - 1;
    }
    run += total;
  }
}

__global__ __launch_bounds__(kBS) void BfsFinish(int32_t* __restrict__ parent, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    if (parent[i] < -1) parent[i] = -1;
}

// Vertices without in-neighbours, once per graph: bottom-up treats them as
// visited (it could never discover them).  Synchronises the stream the
// first time.
uint64_t* NoInNeighbours(gg_graph* g) {
  const int64_t n = g->n;
  const int64_t nwords = (n + 63) / 64;
  hipStream_t st = g->stream;
  const dim3 blk(kBS);
  const dim3 grid_few(GridFor(nwords, kBS / kWave, g->cu_count));
  uint64_t* no_in = g->Scratch<uint64_t>("graph.no_in", nwords);
  if (g->no_in_count < 0) {
    unsigned long long* cnt = g->Scratch<unsigned long long>("graph.no_in_count", 1);
    GG_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st));
    BfsNoInNeighbours<<<grid_few, blk, 0, st>>>(g->in_off, n, no_in, cnt);
    unsigned long long h = 0;
    GG_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    g->no_in_count = static_cast<int64_t>(h);
  }

  return no_in;
}

}  // namespace

// DOBFS from `source` (a device id) on g->stream, inside the caller's timed
// region: parent in the "bfs.parent" workspace array, depth (device, may be
// null) the level of every vertex (-1 unreached).
BfsRun RunDobfs(gg_graph* g, int32_t source, int alpha, int beta, int32_t* depth) {
  const int64_t n = g->n;
  const int64_t nwords = (n + 63) / 64;
  int32_t* parent = g->Scratch<int32_t>("bfs.parent", n);
  int32_t* qin = g->Scratch<int32_t>("bfs.qa", n);
  int32_t* qout = g->Scratch<int32_t>("bfs.qb", n);
  uint64_t* front = g->Scratch<uint64_t>("bfs.front", nwords);
  uint64_t* next = g->Scratch<uint64_t>("bfs.next", nwords);
  uint64_t* visited = g->Scratch<uint64_t>("bfs.visited", nwords);
  BfsCounters* dc = g->Scratch<BfsCounters>("bfs.ctr", 1);
  const bool use_chunks = g->max_degree >= kExpandBS;
  EdgeChunk* chunks =
      use_chunks ? g->Scratch<EdgeChunk>("bfs.chunks", MaxChunks(n, g->out_entries)) : nullptr;

  const dim3 blk(kBS);
  const dim3 grid_n(GridFor(n, kBS, g->cu_count * 32));
  const dim3 grid_w(GridFor(nwords, kBS / kWave, g->cu_count * 32));
  const dim3 grid_few(GridFor(nwords, kBS / kWave, g->cu_count));
  const double N = static_cast<double>(n), W = static_cast<double>(nwords);
  int64_t launches = 0;
  unsigned long long mail[kMailValues];
  auto sum = [&](int k) { return static_cast<int64_t>(mail[k]); };
  // Posts this step's counter totals to the host (and clears them).
  auto publish = [&]() {
    const unsigned long long seq = MailNext(g);
    Run(g, "bfs_publish", 0, BfsPublish, dim3(1), blk, dc, g->mail_dev, seq);
    launches++;
    MailWait(g, seq, mail, kMailValues);
  };

  uint64_t* no_in = NoInNeighbours(g);
  Run(g, "bfs_reset", 0, BfsResetCounters, dim3(1), blk, dc);
  // Algorithmic bytes per launch: the compulsory traffic of each kernel
  // (offsets 8 B, ids 4 B, per-vertex state 4 B, bitmaps N/8).
  Run(g, "bfs_init", 8 * (N + 1) + 4 * N + (depth ? 4 * N : 0), BfsInit, grid_n, blk,
      g->out_off, g->in_off, parent, depth, n, source, qin, dc);
  launches += 2;
  publish();
  int64_t scout = sum(kScout);
  int64_t reach = sum(kReach);
  // Unvisited vertices without in-neighbours that the reference's
  // bottom-up step examines (with zero edges) and this one skips.
  const int64_t skipped_no_in = g->no_in_count - (mail[kMailSrcNoIn] ? 1 : 0);
  int64_t examined = 0, vexamined = 0, reached = 1;
  const int64_t e_dir = g->directed ? g->out_entries : 2 * (g->out_entries / 2);
  int64_t edges_to_check = e_dir;
  int64_t qsize = 1;
  int32_t level = 1;
  int64_t td = 0, bu = 0;
  // Upper bound on the summed out-degree of the current frontier.
  int64_t frontier_deg = scout;
  while (qsize > 0) {
    if (scout > edges_to_check / alpha) {
      Run(g, "bfs_bitmap_clear", 8 * W, FillU64, grid_w, blk, front, nwords, uint64_t(0));
      Run(g, "bfs_queue_to_bitmap", 4.0 * qsize, BfsQueueToBitmap,
          dim3(GridFor(qsize, kBS, g->cu_count * 32)), blk, qin, static_cast<int32_t>(qsize),
          reinterpret_cast<unsigned long long*>(front));
      Run(g, "bfs_visited_from_parent", 4 * N + 16 * W, BfsVisitedFromParent, grid_w, blk, parent, n,
          no_in, visited);
      launches += 3;
      int64_t awake = qsize, old_awake;
      do {
        old_awake = awake;
        const int idx = Run(g, "bfs_bottom_up", 3 * W * 8, BfsBottomUp, grid_w, blk, g->in_off,
                            g->in_dst, g->out_off, parent, depth, front, next, visited, n, dc, level);
        launches++;
        publish();
        awake = sum(kAwake);
        frontier_deg = sum(kReach);
        reached += awake;
        reach += frontier_deg;
        examined += sum(kExamined);
        vexamined += sum(kVexam) + skipped_no_in;
        CreditBytes(g, idx, 4.0 * sum(kExamined) + 16.0 * sum(kVexam));
        level++;
        bu++;
        std::swap(front, next);
      } while (awake >= old_awake || awake > n / beta);
      Run(g, "bfs_bitmap_to_queue", 16 * W + 4.0 * awake, BfsBitmapToQueue, grid_few, blk, front,
          nwords, qin, &dc->bmtail);
      launches++;
      qsize = awake;  // the queue holds exactly the last step's discoveries
      scout = 1;
    } else {
      edges_to_check -= scout;
      const bool hubs = use_chunks && frontier_deg >= kExpandBS;
      // A small frontier gets every vertex with >= 256 edges chunked across
      // the grid; a large one only the real hubs.
      const int64_t chunk_min = qsize <= 4096 ? kExpandBS : kChunkEdges;
      const int idx = Run(g, "bfs_top_down", 16.0 * qsize, BfsTopDown,
                          dim3(GridFor(qsize, kBS, g->cu_count * 8)), blk, g->out_off, g->out_dst,
                          parent, depth, qin, static_cast<int32_t>(qsize), qout, dc, level,
                          hubs ? chunks : static_cast<EdgeChunk*>(nullptr), chunk_min);
      launches++;
      if (hubs) {
        const dim3 grid_c(GridFor(frontier_deg / kExpandBS + kChunkSlices, 1, g->cu_count * 8));
        Run(g, "bfs_top_down_chunks", 0, BfsTopDownChunks, grid_c, blk, g->out_off, g->out_dst,
            parent, depth, qout, dc, level, static_cast<const EdgeChunk*>(chunks));
        launches++;
      }
      vexamined += qsize;
      publish();
      qsize = static_cast<int64_t>(mail[kMailQtail]);
      reached += qsize;
      scout = sum(kScout);
      frontier_deg = scout;
      reach += sum(kReach);
      examined += sum(kExamined);
      CreditBytes(g, idx, 4.0 * sum(kExamined));
      level++;
      td++;
      std::swap(qin, qout);
    }
  }
  Run(g, "bfs_finish", 4 * N, BfsFinish, grid_n, blk, parent, n);
  launches++;
  GG_HIP(hipGetLastError());
  BfsRun r;
  r.td = td;
  r.bu = bu;
  r.levels = level;
  r.reached = reached;
  r.reach = reach;
  r.examined = examined;
  r.vexamined = vexamined;
  r.launches = launches;
  return r;
}

}  // namespace gg

using namespace gg;

extern "C" int gg_bfs(gg_graph* g, int32_t source, int alpha, int beta, int32_t* parent_out,
                      int32_t* depth_out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "BFS needs an unweighted graph upload");
    GG_REQUIRE(source >= 0 && source < g->n, GG_EINVAL, "source out of range");
    GG_REQUIRE(alpha > 0 && beta > 0, GG_EINVAL, "alpha and beta must be positive");
    GG_HIP(hipSetDevice(g->device));
    source = DeviceVertex(g, source);
    int32_t* depth = depth_out ? g->Scratch<int32_t>("bfs.depth", g->n) : nullptr;
    NoInNeighbours(g);
    BeginTiming(g);
    const BfsRun r = RunDobfs(g, source, alpha, beta, depth);
    FinishTiming(g, stats, wt);
    if (parent_out) CopyOutVertexI32(g, g->Scratch<int32_t>("bfs.parent", g->n), parent_out, true);
    if (depth_out) CopyOutVertexI32(g, depth, depth_out, false);
    if (stats) {
      stats->total_ms = wt.ms();
      stats->iterations = r.td + r.bu;
      stats->td_steps = r.td;
      stats->bu_steps = r.bu;
      stats->reached = r.reached;
      stats->reached_edges = r.reach;
      stats->edges_examined = r.examined;
      stats->vertices_examined = r.vexamined;
      stats->launches = r.launches;
    }
  });
}
