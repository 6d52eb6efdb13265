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
//   * the top-down step is one kernel that expands each 256-vertex tile of
//     the frontier three ways: vertices with >= 256 edges by the whole block,
//     >= 64 by one wavefront, the rest by a block-wide prefix sum over their
//     edges staged in LDS; discovered vertices go to the next queue with one
//     atomic per wavefront (ballot + popcount);
//   * the bottom-up step gives each wavefront 64 consecutive vertices, so
//     the frontier bitmap word for them is one ballot, written without
//     atomics;
//   * the controller runs on the host and reads back one small counter
//     block per level.
#include "gg_common.h"

namespace gg {
namespace {

constexpr int kBS = 256;

struct BfsCounters {
  int32_t qtail;
  int32_t pad;
  unsigned long long scout;     // TD: sum of -parent of discovered vertices
  unsigned long long awake;     // BU: vertices discovered
  unsigned long long reach;     // sum of out-degrees of discovered vertices
  unsigned long long examined;  // adjacency entries read
  unsigned long long vexam;     // BU: unvisited vertices scanned
};

__global__ void BfsResetCounters(BfsCounters* c) {
  if (threadIdx.x == 0) {
    c->qtail = 0;
    c->scout = 0;
    c->awake = 0;
    c->reach = 0;
    c->examined = 0;
    c->vexam = 0;
  }
}

__global__ __launch_bounds__(kBS) void BfsInit(const int64_t* __restrict__ off, int32_t* __restrict__ parent,
                                               int32_t* __restrict__ depth, int64_t n, int32_t src,
                                               int32_t* __restrict__ queue, BfsCounters* c) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    const int64_t deg = off[i + 1] - off[i];
    int32_t p = deg != 0 ? static_cast<int32_t>(-deg) : -1;
    if (i == src) {
      p = src;
      queue[0] = src;
      c->scout = static_cast<unsigned long long>(deg);
      c->reach = static_cast<unsigned long long>(deg);
    }
    parent[i] = p;
    if (depth) depth[i] = (i == src) ? 0 : -1;
  }
}

// One frontier edge u -> v (valid lanes only act; all lanes must call).
struct TdVisit {
  int32_t* parent;
  int32_t* depth;
  const int64_t* off;
  int32_t* qout;
  int32_t* qtail;
  int32_t level;
  unsigned long long scout = 0, reach = 0;

  __device__ __forceinline__ void operator()(int32_t u, int32_t v, bool valid) {
    bool won = false;
    int32_t cur = 0;
    if (valid) {
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
    WaveAppend(won, v, qout, qtail);
  }
};

__global__ __launch_bounds__(kBS) void BfsTopDown(const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ dst,
                                                  int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                                  const int32_t* __restrict__ qin, int32_t qn,
                                                  int32_t* __restrict__ qout, BfsCounters* c, int32_t level) {
  __shared__ int s_claim;
  __shared__ int32_t s_u;
  __shared__ int64_t s_beg, s_end;
  __shared__ int s_scan[kBS / kWave];
  __shared__ int32_t s_src[kBS];
  __shared__ int64_t s_eidx[kBS];
  __shared__ unsigned long long s_red[kBS / kWave];

  const int tid = threadIdx.x, lane = LaneId();
  TdVisit visit{parent, depth, off, qout, &c->qtail, level};
  unsigned long long examined = 0;

  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < qn; tile += int64_t(gridDim.x) * kBS) {
    const int64_t i = tile + tid;
    int32_t u = 0;
    int64_t beg = 0, end = 0;
    if (i < qn) {
      u = qin[i];
      beg = off[u];
      end = off[u + 1];
      examined += static_cast<unsigned long long>(end - beg);
    }
    // Vertices with >= kBS edges: the whole block, one vertex at a time.
    while (true) {
      if (tid == 0) s_claim = -1;
      __syncthreads();
      if (end - beg >= kBS) s_claim = tid;
      __syncthreads();
      const int owner = s_claim;
      if (owner < 0) break;
      if (tid == owner) {
        s_u = u;
        s_beg = beg;
        s_end = end;
        beg = end;
      }
      __syncthreads();
      const int32_t bu = s_u;
      const int64_t bb = s_beg, be = s_end;
      for (int64_t base = bb; base < be; base += kBS) {
        const int64_t e = base + tid;
        const bool ok = e < be;
        visit(bu, ok ? dst[e] : 0, ok);
      }
      __syncthreads();
    }
    // Vertices with >= 64 edges: one wavefront each.
    while (true) {
      const uint64_t m = __ballot(end - beg >= kWave);
      if (m == 0) break;
      const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
      const int32_t wu = __shfl(u, leader);
      const int64_t wb = __shfl(beg, leader), we = __shfl(end, leader);
      if (lane == leader) beg = end;
      for (int64_t base = wb; base < we; base += kWave) {
        const int64_t e = base + lane;
        const bool ok = e < we;
        visit(wu, ok ? dst[e] : 0, ok);
      }
    }
    // The rest (< 64 edges each): block-wide prefix sum over their edges.
    const int d = static_cast<int>(end - beg);
    int total = 0;
    const int rank = BlockExclusiveScan<kBS>(d, &total, s_scan);
    for (int done = 0; done < total; done += kBS) {
      const int lo = rank > done ? rank : done;
      const int hi = (rank + d) < (done + kBS) ? (rank + d) : (done + kBS);
      for (int j = lo; j < hi; j++) {
        s_src[j - done] = u;
        s_eidx[j - done] = beg + (j - rank);
      }
      __syncthreads();
      const bool ok = done + tid < total;
      const int32_t su = ok ? s_src[tid] : 0;
      const int32_t v = ok ? dst[s_eidx[tid]] : 0;
      visit(su, v, ok);
      __syncthreads();
    }
  }
  const unsigned long long sc = BlockSum<kBS>(visit.scout, s_red);
  const unsigned long long re = BlockSum<kBS>(visit.reach, s_red);
  const unsigned long long ex = BlockSum<kBS>(examined, s_red);
  if (tid == 0) {
    if (sc) atomicAdd(&c->scout, sc);
    if (re) atomicAdd(&c->reach, re);
    if (ex) atomicAdd(&c->examined, ex);
  }
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

// Each wavefront owns bitmap words; lane l handles vertex 64*word + l.
__global__ __launch_bounds__(kBS) void BfsBottomUp(const int64_t* __restrict__ in_off,
                                                   const int32_t* __restrict__ in_dst,
                                                   const int64_t* __restrict__ out_off,
                                                   int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                                   const uint64_t* __restrict__ front,
                                                   uint64_t* __restrict__ next, int64_t n, BfsCounters* c,
                                                   int32_t level) {
  __shared__ unsigned long long s_red[kBS / kWave];
  const int lane = LaneId();
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  unsigned long long awake = 0, reach = 0, examined = 0, vexam = 0;
  for (int64_t w = wave; w < nwords; w += nwaves) {
    const int64_t u = w * 64 + lane;
    bool found = false;
    if (u < n && parent[u] < 0) {
      vexam++;
      const int64_t b = in_off[u], e = in_off[u + 1];
      int64_t k = b;
      for (; k < e; k++) {
        const int32_t v = in_dst[k];
        if ((front[v >> 6] >> (v & 63)) & 1) {
          parent[u] = v;
          found = true;
          break;
        }
      }
      examined += static_cast<unsigned long long>((found ? k + 1 : e) - b);
      if (found) {
        if (depth) depth[u] = level;
        reach += static_cast<unsigned long long>(out_off[u + 1] - out_off[u]);
      }
    }
    const uint64_t m = __ballot(found);
    if (lane == 0) {
      next[w] = m;
      awake += __popcll(m);
    }
  }
  const unsigned long long aw = BlockSum<kBS>(awake, s_red);
  const unsigned long long re = BlockSum<kBS>(reach, s_red);
  const unsigned long long ex = BlockSum<kBS>(examined, s_red);
  const unsigned long long vx = BlockSum<kBS>(vexam, s_red);
  if (threadIdx.x == 0) {
    if (aw) atomicAdd(&c->awake, aw);
    if (re) atomicAdd(&c->reach, re);
    if (ex) atomicAdd(&c->examined, ex);
    if (vx) atomicAdd(&c->vexam, vx);
  }
}

__global__ __launch_bounds__(kBS) void BfsBitmapToQueue(const uint64_t* __restrict__ bm, int64_t n,
                                                        int32_t* __restrict__ q, BfsCounters* c) {
  const int lane = LaneId();
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  for (int64_t w = wave; w < nwords; w += nwaves) {
    const bool set = (bm[w] >> lane) & 1;
    WaveAppend(set, static_cast<int32_t>(w * 64 + lane), q, &c->qtail);
  }
}

__global__ __launch_bounds__(kBS) void BfsFinish(int32_t* __restrict__ parent, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    if (parent[i] < -1) parent[i] = -1;
}

}  // namespace
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
    const int64_t n = g->n;
    const int64_t nwords = (n + 63) / 64;
    hipStream_t st = g->stream;

    int32_t* parent = g->Scratch<int32_t>("bfs.parent", n);
    int32_t* depth = depth_out ? g->Scratch<int32_t>("bfs.depth", n) : nullptr;
    int32_t* qin = g->Scratch<int32_t>("bfs.qa", n);
    int32_t* qout = g->Scratch<int32_t>("bfs.qb", n);
    uint64_t* front = g->Scratch<uint64_t>("bfs.front", nwords);
    uint64_t* next = g->Scratch<uint64_t>("bfs.next", nwords);
    BfsCounters* dc = g->Scratch<BfsCounters>("bfs.ctr", 1);
    BfsCounters* hc = g->Pinned<BfsCounters>(1);

    const unsigned grid_n = GridFor(n, kBS, g->cu_count * 32);
    const unsigned grid_w = GridFor(nwords, kBS / kWave, g->cu_count * 32);
    const double N = static_cast<double>(n), W = static_cast<double>(nwords);
    int64_t launches = 0;
    auto read_counters = [&]() {
      GG_HIP(hipMemcpyAsync(hc, dc, sizeof(BfsCounters), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
    };
    auto reset = [&]() {
      BfsResetCounters<<<1, 64, 0, st>>>(dc);
      launches++;
    };

    BeginTiming(g);
    reset();
    // Algorithmic bytes per launch: the compulsory traffic of each kernel
    // (offsets 8 B, ids 4 B, per-vertex state 4 B, bitmaps N/8).
    Launch(g, "bfs_init", 8 * (N + 1) + 4 * N + (depth ? 4 * N : 0), [&] {
      BfsInit<<<grid_n, kBS, 0, st>>>(g->out_off, parent, depth, n, source, qin, dc);
    });
    launches++;
    read_counters();
    int64_t scout = static_cast<int64_t>(hc->scout);
    int64_t reach = static_cast<int64_t>(hc->reach);
    int64_t examined = 0, vexamined = 0, reached = 1;
    const int64_t e_dir = g->directed ? g->out_entries : 2 * (g->out_entries / 2);
    int64_t edges_to_check = e_dir;
    int64_t qsize = 1;
    int32_t level = 1;
    int64_t td = 0, bu = 0;
    while (qsize > 0) {
      if (scout > edges_to_check / alpha) {
        Launch(g, "bfs_bitmap_clear", 8 * W,
               [&] { FillU64<<<grid_w, kBS, 0, st>>>(front, nwords, 0); });
        Launch(g, "bfs_queue_to_bitmap", 4.0 * qsize, [&] {
          BfsQueueToBitmap<<<GridFor(qsize, kBS, g->cu_count * 32), kBS, 0, st>>>(
              qin, static_cast<int32_t>(qsize), reinterpret_cast<unsigned long long*>(front));
        });
        launches += 2;
        int64_t awake = qsize, old_awake;
        do {
          old_awake = awake;
          reset();
          const int idx = Launch(g, "bfs_bottom_up", 2 * W * 8, [&] {
            BfsBottomUp<<<grid_w, kBS, 0, st>>>(g->in_off, g->in_dst, g->out_off, parent, depth,
                                                front, next, n, dc, level);
          });
          launches++;
          read_counters();
          awake = static_cast<int64_t>(hc->awake);
          reached += awake;
          reach += static_cast<int64_t>(hc->reach);
          examined += static_cast<int64_t>(hc->examined);
          vexamined += static_cast<int64_t>(hc->vexam);
          CreditBytes(g, idx, 4.0 * hc->examined + 16.0 * hc->vexam);
          level++;
          bu++;
          std::swap(front, next);
        } while (awake >= old_awake || awake > n / beta);
        reset();
        Launch(g, "bfs_bitmap_to_queue", 8 * W + 4.0 * awake,
               [&] { BfsBitmapToQueue<<<grid_w, kBS, 0, st>>>(front, n, qin, dc); });
        launches++;
        qsize = awake;  // the queue holds exactly the last step's discoveries
        scout = 1;
      } else {
        edges_to_check -= scout;
        reset();
        const int idx = Launch(g, "bfs_top_down", 16.0 * qsize, [&] {
          BfsTopDown<<<GridFor(qsize, kBS, g->cu_count * 16), kBS, 0, st>>>(
              g->out_off, g->out_dst, parent, depth, qin, static_cast<int32_t>(qsize), qout, dc,
              level);
        });
        launches++;
        vexamined += qsize;
        read_counters();
        qsize = hc->qtail;
        reached += qsize;
        scout = static_cast<int64_t>(hc->scout);
        reach += static_cast<int64_t>(hc->reach);
        examined += static_cast<int64_t>(hc->examined);
        CreditBytes(g, idx, 4.0 * hc->examined);
        level++;
        td++;
        std::swap(qin, qout);
      }
    }
    Launch(g, "bfs_finish", 4 * N, [&] { BfsFinish<<<grid_n, kBS, 0, st>>>(parent, n); });
    launches++;
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (parent_out) GG_HIP(hipMemcpy(parent_out, parent, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (depth_out) GG_HIP(hipMemcpy(depth_out, depth, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (stats) {
      stats->total_ms = wt.ms();
      stats->iterations = td + bu;
      stats->td_steps = td;
      stats->bu_steps = bu;
      stats->reached = reached;
      stats->reached_edges = reach;
      stats->edges_examined = examined;
      stats->vertices_examined = vexamined;
      stats->launches = launches;
    }
  });
}
