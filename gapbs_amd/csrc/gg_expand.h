// Load-balanced expansion of a vertex frontier over its CSR edges.
//
// One 256-thread block takes a tile of 256 frontier vertices and visits all
// their edges so that a hub with millions of edges and a vertex with three
// cost about the same per edge:
//   0. vertices with >= kChunkEdges edges are cut into chunks of
//      kChunkEdges edges appended to a chunk list, which a second kernel
//      (ExpandChunks) spreads over the whole grid;
//   1. vertices with >= 256 edges: the whole block, one vertex at a time;
//   2. vertices with >= 64 edges: one wavefront each;
//   3. the rest: a block-wide prefix sum over their edge counts, the edge
//      indices staged in LDS, one edge per thread.
// The visitor is called as op(u, payload, edge_index, valid) by every lane
// of a wavefront together (valid == false for padding lanes), so it may use
// wavefront-wide ballots (e.g. WaveAppend).
#pragma once

#include "gg_common.h"

namespace gg {

constexpr int kExpandBS = 256;
constexpr int64_t kChunkEdges = 4096;

struct EdgeChunk {
  int64_t beg, end;
  int32_t u, payload;
};

// Upper bound on the chunks one expansion can emit when vertices with at
// least kExpandBS edges are chunked: every frontier vertex appears once, so
// sum(ceil(deg / kChunkEdges)) <= E / kChunkEdges + #{deg >= kExpandBS}.
inline int64_t MaxChunks(int64_t n, int64_t entries) {
  const int64_t big = entries / kExpandBS < n ? entries / kExpandBS : n;
  return entries / kChunkEdges + big + 2;
}

struct ExpandShared {
  int claim;
  int32_t u;
  int32_t payload;
  int64_t beg, end;
  int scan[kExpandBS / kWave];
  int32_t src[kExpandBS];
  int32_t pay[kExpandBS];
  int64_t eidx[kExpandBS];
};

// Expands one tile.  (u, payload, beg, end) describe this thread's vertex
// (beg == end for threads without one).  Must be called by all threads of
// the block.  Vertices with at least chunk_min (>= kExpandBS) edges are
// handed to the chunk list; chunks == nullptr disables that step.
template <class Op>
__device__ __forceinline__ void ExpandTile(ExpandShared& sh, int32_t u, int32_t payload,
                                           int64_t beg, int64_t end, Op& op,
                                           EdgeChunk* chunks, int32_t* nchunks, int64_t chunk_min) {
  const int tid = threadIdx.x, lane = LaneId();
  // 0. Hubs: hand their edges to the grid-wide chunk kernel.
  if (chunks && end - beg >= chunk_min) {
    const int64_t d = end - beg;
    const int nc = static_cast<int>((d + kChunkEdges - 1) / kChunkEdges);
    const int base = atomicAdd(nchunks, nc);
    for (int c = 0; c < nc; c++) {
      const int64_t cb = beg + c * kChunkEdges;
      const int64_t ce = cb + kChunkEdges < end ? cb + kChunkEdges : end;
      chunks[base + c] = EdgeChunk{cb, ce, u, payload};
    }
    beg = end;
  }
  // 1. >= kExpandBS edges: whole block.
  while (true) {
    if (tid == 0) sh.claim = -1;
    __syncthreads();
    if (end - beg >= kExpandBS) sh.claim = tid;
    __syncthreads();
    const int owner = sh.claim;
    if (owner < 0) break;
    if (tid == owner) {
      sh.u = u;
      sh.payload = payload;
      sh.beg = beg;
      sh.end = end;
      beg = end;
    }
    __syncthreads();
    const int32_t bu = sh.u, bp = sh.payload;
    const int64_t bb = sh.beg, be = sh.end;
    for (int64_t base = bb; base < be; base += kExpandBS) {
      const int64_t e = base + tid;
      op(bu, bp, e, e < be);
    }
    __syncthreads();
  }
  // 2. >= 64 edges: one wavefront.
  while (true) {
    const uint64_t m = __ballot(end - beg >= kWave);
    if (m == 0) break;
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    const int32_t wu = __shfl(u, leader), wp = __shfl(payload, leader);
    const int64_t wb = __shfl(beg, leader), we = __shfl(end, leader);
    if (lane == leader) beg = end;
    for (int64_t base = wb; base < we; base += kWave) {
      const int64_t e = base + lane;
      op(wu, wp, e, e < we);
    }
  }
  // 3. < 64 edges each: block-wide prefix sum, edges staged in LDS.
  const int d = static_cast<int>(end - beg);
  int total = 0;
  const int rank = BlockExclusiveScan<kExpandBS>(d, &total, sh.scan);
  for (int done = 0; done < total; done += kExpandBS) {
    const int lo = rank > done ? rank : done;
    const int hi = (rank + d) < (done + kExpandBS) ? (rank + d) : (done + kExpandBS);
    for (int j = lo; j < hi; j++) {
      sh.src[j - done] = u;
      sh.pay[j - done] = payload;
      sh.eidx[j - done] = beg + (j - rank);
    }
    __syncthreads();
    const bool ok = done + tid < total;
    op(ok ? sh.src[tid] : 0, ok ? sh.pay[tid] : 0, ok ? sh.eidx[tid] : 0, ok);
    __syncthreads();
  }
}

// Visits the edges of every chunk emitted by ExpandTile.  Work items are
// 256-edge slices of chunks, spread over the grid.  Must be called by all
// threads of the block.
constexpr int kChunkSlices = static_cast<int>(kChunkEdges / kExpandBS);

template <class Op>
__device__ __forceinline__ void ExpandChunks(const EdgeChunk* __restrict__ chunks, int32_t nchunks,
                                             Op& op) {
  const int64_t items = int64_t(nchunks) * kChunkSlices;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const EdgeChunk ch = chunks[it / kChunkSlices];
    const int64_t e = ch.beg + (it % kChunkSlices) * kExpandBS + threadIdx.x;
    op(ch.u, ch.payload, e, e < ch.end);
  }
}

// A block's output queue staged in LDS: wavefronts append with LDS atomics,
// the block flushes once with one global atomic.  Entries that do not fit
// go straight to global memory.
template <int CAP>
struct LdsQueue {
  int count;  // entries reserved; may exceed CAP
  int32_t base;
  int32_t items[CAP];
};

template <int CAP>
__device__ __forceinline__ void LdsQueueInit(LdsQueue<CAP>& q) {
  if (threadIdx.x == 0) q.count = 0;
  __syncthreads();
}

// Wave-wide: every active lane of the wavefront calls it together.
template <int CAP>
__device__ __forceinline__ void LdsQueuePush(LdsQueue<CAP>& q, bool take, int32_t v, int32_t* gout,
                                             int32_t* gtail) {
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const int lane = LaneId();
  const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  const int cnt = __popcll(m);
  int base = 0;
  if (lane == leader) base = atomicAdd(&q.count, cnt);
  base = __shfl(base, leader);
  const int fit = CAP - base > 0 ? CAP - base : 0;
  const int over = cnt > fit ? cnt - fit : 0;
  int gbase = 0;
  if (over && lane == leader) gbase = atomicAdd(gtail, over);
  gbase = __shfl(gbase, leader);
  if (take) {
    const int rank = __popcll(LanesBelow(m));
    if (rank < fit)
      q.items[base + rank] = v;
    else
      gout[gbase + rank - fit] = v;
  }
}

// Block-wide: all threads call it (synchronises).
template <int CAP>
__device__ __forceinline__ void LdsQueueFlush(LdsQueue<CAP>& q, int32_t* gout, int32_t* gtail) {
  __syncthreads();
  const int cnt = q.count < CAP ? q.count : CAP;
  if (threadIdx.x == 0) q.base = cnt ? atomicAdd(gtail, cnt) : 0;
  __syncthreads();
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) gout[q.base + i] = q.items[i];
}

// Block-aggregated append: one atomic per block.  All threads of the block
// call it together; `sh` needs kExpandBS/64 + 1 ints.
template <class T>
__device__ __forceinline__ void BlockAppend(bool take, T value, T* out, int32_t* tail, int* sh) {
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  const uint64_t m = __ballot(take);
  if (lane == 0) sh[wid] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < kExpandBS / kWave; w++) {
      const int c = sh[w];
      sh[w] = total;
      total += c;
    }
    sh[kExpandBS / kWave] = total ? atomicAdd(tail, total) : 0;
  }
  __syncthreads();
  if (take) out[sh[kExpandBS / kWave] + sh[wid] + __popcll(LanesBelow(m))] = value;
  __syncthreads();
}

}  // namespace gg
