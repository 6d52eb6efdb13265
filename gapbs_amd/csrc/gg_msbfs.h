// Batched breadth-first search: up to 64 roots in one traversal (kernels
// and the host level loop, shared by the batched BFS and batched BC entry
// points).
//
// Reference counterpart: none directly -- the reference runs one root per
// trial (src/benchmark.h:106, src/bfs.cc:256-259).  Each root's depths must
// equal what DOBFS gives for it alone (bfs.cc:122-179); depths are unique,
// so this is checked bit for bit.
//
// Design (the multi-source BFS of Then et al., "The More the Merrier", VLDB
// 2015, on wave64): per vertex a 64-bit `seen` mask (bit r: root r reached
// it) and a 64-bit frontier mask (the bits that arrived at the last level).
// A level is either
//   top-down  -- the active vertices' edges are expanded (gg_expand.h) and
//                each neighbour's accumulator gets frontier & ~seen with
//                atomicOr; the first toucher queues the neighbour, and an
//                update pass turns accumulators into new bits; or
//   bottom-up -- every vertex not yet reached by all roots ORs its
//                in-neighbours' frontier masks, stopping once every root has
//                reached it (one thread per vertex below 64 in-edges, one
//                wavefront per heavier vertex from a list built once per
//                graph); no atomics, and the next frontier array is written
//                densely, so no list is built.
// Direction: bottom-up while the active vertices' out-edges exceed
// unfinished / alpha, the form of the reference's test (bfs.cc:143) with
// "unexplored edges" read as the in-edges of vertices that some root has
// not reached yet (the reference's edges_to_check would go negative here:
// a vertex is active at one level per distinct depth).  A top-down level
// after a bottom-up one first compacts the dense frontier into a list.
// Depths are kept vertex-major as uint16 (64 per vertex, 128 bytes),
// written per level in whole 16-byte chunks by a wavefront together; a
// slot is defined only where the vertex's seen mask has the root's bit, so
// the array needs no initialisation.
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
// Grid caps of the level kernels, in blocks per CU (top-down 1/2 and 1/4
// of it): 64 instead of 16 (bottom-up tiers) and 8 / 4 (top-down) took the
// batched BFS of 64 roots at Kronecker 26 from 31.3 to 30.2 ms.
constexpr unsigned kLevelBlocksPerCu = 64;
constexpr int kQueueCap = 2048;
constexpr int kStripes = 64;
constexpr uint16_t kNoDepth = 0xFFFF;
// Bottom-up: light tier below kHeavyEdges in-edges, heavy tier from there.  Kronecker 26, 64
// roots (heavy chunks ordered by first source): 512 -> 27.5 ms, 1024 -> 26.2 ms, 2048 -> 25.9 ms.
constexpr int64_t kHeavyEdges = 2048;
constexpr int64_t kHeavyChunk = 2048;  // heavy in-edges per wavefront work item
constexpr int kUnroll = 4;             // gathers in flight per lane
// Light tier: list positions scanned per round (a vertex leaves once it
// has every bit it can still gain).
constexpr int kRounds = 3;
__constant__ const int kRoundEnd[kRounds] = {8, 64, static_cast<int>(kHeavyEdges)};  // {16, 64}: 2% slower

struct MsCounters {
  int32_t active_tail;   // top-down: next level's active list
  int32_t touched_tail;  // top-down: vertices whose accumulator was touched
  int32_t compact_tail;  // dense frontier -> list
  int32_t nchunks;
  int32_t heavy_tail;    // heavy-vertex list (built once per graph)
  int32_t heavy_chunks;  // and its chunks
  unsigned long long awake[kStripes];     // bottom-up: vertices with new bits
  unsigned long long scout[kStripes];     // out-degree of vertices with new bits
  unsigned long long examined[kStripes];  // adjacency entries read (top-down, light bottom-up)
  unsigned long long heavy_ex[kStripes];  // adjacency entries read by the heavy tier
  unsigned long long blk_ex[kStripes];    // in-entries read by the propagation-block pass
  unsigned long long finished[kStripes];  // in-edges of vertices that became reached by all roots
  unsigned long long live[kStripes];      // OR of the new bits
  // Depth-row bytes written (and read first) by the three kernels that
  // write depths: [0] the bottom-up light tier, [1] the top-down update,
  // [2] the heavy tier's finish (for the profile's byte credits).
  unsigned long long depth_bytes[3][kStripes];
};

// Level state, kept on the device and advanced by MsStep after every level,
// so no level kernel needs a decision from the host: each reads the state
// and returns at once unless its direction runs now.  This lets a run of
// levels be replayed from a captured graph without host round trips.
struct MsLevel {
  int64_t na;          // vertices with bits in the current frontier
  int64_t na_prev;     // the same for the frontier the last level consumed
  int64_t unfinished;  // in-edges of vertices some root has not reached
  int64_t td, bu;      // levels run in each direction
  int64_t examined;    // adjacency entries read so far
  uint64_t live;       // roots with bits in the current frontier
  int32_t level;       // the level about to run (1-based)
  int32_t bottom_up;   // its direction
  int32_t listed;      // the current frontier is listed in the active list
  int32_t clear;       // to clear in the array just consumed: 0 nothing, 1 its listed entries, 2 all
  int32_t done;        // no frontier left
  int32_t blocks;      // this (bottom-up) level goes through the propagation blocks
  int64_t blocks_min;  // ... (the first one after a top-down level) when more in-edges than this are unfinished
};

// Mail posted by MsStep for the host-driven loop.
enum {
  kMailNa, kMailBottomUp, kMailDone, kMailExamined, kMailHeavyEx, kMailLevel, kMailBlocksEx,
  kMailDepthBottomUp, kMailDepthUpdate, kMailDepthHeavy, kMailValues
};

__device__ __forceinline__ uint64_t WaveOr(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// Block-wide OR, result valid in every thread.  `sh` needs BS/64 slots.
template <int BS>
__device__ __forceinline__ unsigned long long BlockOr(unsigned long long v, unsigned long long* sh) {
  v = WaveOr(v);
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  unsigned long long t = 0;
#pragma unroll
  for (int w = 0; w < BS / kWave; w++) t |= sh[w];
  __syncthreads();
  return t;
}

// Starts the level state after the roots (their distinct vertices are the
// first active list).
__global__ void MsLevelInit(MsCounters* c, MsLevel* L, int64_t unfinished, uint64_t full, int64_t blocks_min) {
  if (threadIdx.x == 0) {
    L->blocks = 0;
    L->blocks_min = blocks_min;
    const int64_t na = c->active_tail;
    c->active_tail = c->touched_tail = c->compact_tail = c->nchunks = 0;
    L->na = na;
    L->na_prev = 0;
    L->unfinished = unfinished;
    L->td = L->bu = L->examined = 0;
    L->live = full;
    L->level = 1;
    L->bottom_up = 0;  // level 1 is top-down whatever the roots' degrees
    L->listed = 1;
    L->clear = 0;
    L->done = na == 0;
  }
}

// After a level: folds the striped counters into the level state and picks
// the next direction -- bottom-up while the new frontier's out-edges exceed
// unfinished / alpha -- and what the next level must clear first (a
// top-down level writes only the entries that gain bits).  With mb, posts
// the state for the host-driven loop.  A no-op once done.
__global__ void MsStep(MsCounters* c, MsLevel* L, int alpha, int keep, unsigned long long* mb,
                       unsigned long long seq) {
  if (threadIdx.x == 0) {
    if (!L->done) {
      unsigned long long a = 0, s = 0, ex = 0, hx = 0, sx = 0, fin = 0, lv = 0, db[3] = {0, 0, 0};
      for (int i = 0; i < kStripes; i++) {
        for (int k = 0; k < 3; k++) {
          db[k] += c->depth_bytes[k][i];
          c->depth_bytes[k][i] = 0;
        }
        a += c->awake[i];
        s += c->scout[i];
        ex += c->examined[i];
        hx += c->heavy_ex[i];
        sx += c->blk_ex[i];
        fin += c->finished[i];
        lv |= c->live[i];
        c->awake[i] = c->scout[i] = c->examined[i] = c->heavy_ex[i] = c->blk_ex[i] = c->finished[i] = c->live[i] = 0;
      }
      const int64_t na_next = static_cast<int64_t>(a) + c->active_tail;
      c->active_tail = c->touched_tail = c->compact_tail = c->nchunks = 0;
      const int was_bu = L->bottom_up;
      if (was_bu)
        L->bu++;
      else
        L->td++;
      L->examined += static_cast<int64_t>(ex + hx + sx);
      L->unfinished -= static_cast<int64_t>(fin);
      const int next_bu = static_cast<int64_t>(s) > L->unfinished / alpha;
      L->blocks = next_bu && !was_bu && L->unfinished > L->blocks_min;
      L->clear = (!keep && na_next > 0 && !next_bu) ? (L->listed ? 1 : 2) : 0;
      L->na_prev = L->na;
      L->listed = !was_bu;
      L->bottom_up = next_bu;
      L->na = na_next;
      L->live = lv;
      L->level++;
      L->done = na_next == 0;
      if (mb) {
        const unsigned long long vals[kMailValues] = {static_cast<unsigned long long>(na_next),
                                                      static_cast<unsigned long long>(next_bu),
                                                      static_cast<unsigned long long>(L->done), ex, hx,
                                                      static_cast<unsigned long long>(L->level), sx, db[0], db[1], db[2]};
        MailPost(mb, vals, kMailValues, seq);
      }
    } else if (mb) {
      const unsigned long long vals[kMailValues] = {0, 0, 1, 0, 0, static_cast<unsigned long long>(L->level), 0, 0, 0, 0};
      MailPost(mb, vals, kMailValues, seq);
    }
  }
}

__global__ __launch_bounds__(kBS) void MsInit(uint64_t* __restrict__ seen, uint64_t* __restrict__ f0,
                                              uint64_t* __restrict__ f1, uint64_t* __restrict__ acc, int64_t n,
                                              MsCounters* c) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    seen[i] = 0;
    f0[i] = 0;
    f1[i] = 0;
    acc[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kStripes) {
    const int t = threadIdx.x;
    c->awake[t] = c->scout[t] = c->examined[t] = c->heavy_ex[t] = c->blk_ex[t] = c->finished[t] = c->live[t] = 0;
    c->depth_bytes[0][t] = c->depth_bytes[1][t] = c->depth_bytes[2][t] = 0;
    if (t == 0) c->active_tail = c->touched_tail = c->compact_tail = c->nchunks = 0;
  }
}

// Vertices with at least kHeavyEdges in-edges (once per graph).
__global__ __launch_bounds__(kBS) void MsHeavyList(const int64_t* __restrict__ in_off, int64_t n,
                                                   int32_t* out, MsCounters* c) {
  __shared__ LdsQueue<kQueueCap> lq;
  LdsQueueInit(lq);
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < n; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    const bool heavy = i < n && in_off[i + 1] - in_off[i] >= kHeavyEdges;
    LdsQueuePush(lq, heavy, static_cast<int32_t>(i), out, &c->heavy_tail);
  }
  LdsQueueFlush(lq, out, &c->heavy_tail);
}

// Roots: seen and frontier bits, depth 0, the first active list (each
// distinct root vertex once: the thread whose bit lands first on it; a
// repeated root sets several bits on it).  Root i takes bit bits[i], or bit
// i when bits is null.
__global__ void MsRoots(const int32_t* __restrict__ roots, const int32_t* __restrict__ bits, int count,
                        uint64_t* seen, uint64_t* frontier, uint16_t* depth, int32_t* active, MsCounters* c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const int32_t v = roots[i];
    const int b = bits ? bits[i] : i;
    atomicOr(reinterpret_cast<unsigned long long*>(&seen[v]), 1ull << b);
    const uint64_t old = atomicOr(reinterpret_cast<unsigned long long*>(&frontier[v]), 1ull << b);
    if (depth) depth[int64_t(v) * 64 + b] = 0;
    if (old == 0) active[atomicAdd(&c->active_tail, 1)] = v;
  }
}

// Records `level` as the depth of the new bits nb of vertex w.  Called by
// every lane of a wavefront together, each with its own (w, nb, prior);
// lanes without a vertex pass nb = 0.  prior is the vertex's seen mask
// before this level.  In step i, lane l serves 16-byte chunk l % 8 (roots
// 8c .. 8c+7) of the row of the vertex held by lane 8i + l / 8, so a step
// covers eight whole 128-byte rows.  A chunk with no earlier bits is
// written without being read: slots of roots that have not reached the
// vertex are undefined, and readers mask every row with seen.
__device__ __forceinline__ void WaveWriteDepths(uint16_t* depth, int32_t w, uint64_t nb, uint64_t prior,
                                                uint16_t level) {
  const int lane = LaneId();
  const int c = lane & 7;
  const uint32_t fill = static_cast<uint32_t>(level) * 0x10001u;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int src = 8 * i + (lane >> 3);
    const uint64_t b = __shfl(nb, src);
    const uint64_t p = __shfl(prior, src);
    const int32_t v = __shfl(w, src);
    const uint32_t bits = static_cast<uint32_t>(b >> (8 * c)) & 0xFFu;
    if (bits) {
      uint4* q = reinterpret_cast<uint4*>(depth + int64_t(v) * 64) + c;
      uint4 x = ((p >> (8 * c)) & 0xFFu) ? *q : make_uint4(~0u, ~0u, ~0u, ~0u);
      uint32_t m[4];
#pragma unroll
      for (int k = 0; k < 4; k++)
        m[k] = (((bits >> (2 * k)) & 1u) ? 0x0000FFFFu : 0u) | (((bits >> (2 * k + 1)) & 1u) ? 0xFFFF0000u : 0u);
      x.x = (x.x & ~m[0]) | (fill & m[0]);
      x.y = (x.y & ~m[1]) | (fill & m[1]);
      x.z = (x.z & ~m[2]) | (fill & m[2]);
      x.w = (x.w & ~m[3]) | (fill & m[3]);
      *q = x;
    }
  }
}

// Bytes WaveWriteDepths moves for one vertex: 16 per 8-root chunk with new
// bits, and 16 more for each such chunk it reads first (one that already
// held depths).
__device__ __forceinline__ unsigned long long DepthBytes(uint64_t nb, uint64_t prior) {
  unsigned long long b = 0;
#pragma unroll
  for (int c = 0; c < 8; c++)
    if ((nb >> (8 * c)) & 0xFFu) b += ((prior >> (8 * c)) & 0xFFu) ? 32 : 16;
  return b;
}

// ---- top-down ---------------------------------------------------------------

struct TdVisit {
  const int32_t* dst;
  const uint64_t* frontier;
  const uint64_t* seen;
  uint64_t* acc;
  LdsQueue<kQueueCap>* lq;
  int32_t* touched;
  int32_t* touched_tail;
  __device__ __forceinline__ void operator()(int32_t v, int32_t, int64_t e, bool valid) {
    bool first = false;
    int32_t w = 0;
    if (valid) {
      w = dst[e];
      const uint64_t bits = frontier[v] & ~seen[w];
      // acc[w] only gains bits during the level, so a plain (possibly
      // stale) read that already holds them makes the atomic redundant.
      if (bits && (bits & ~acc[w])) {
        const uint64_t old = atomicOr(reinterpret_cast<unsigned long long*>(&acc[w]), bits);
        first = old == 0;
      }
    }
    LdsQueuePush(*lq, first, w, touched, touched_tail);
  }
};

__global__ __launch_bounds__(kBS) void MsTopDown(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                 const uint64_t* __restrict__ frontier,
                                                 const uint64_t* __restrict__ seen, uint64_t* acc,
                                                 const int32_t* __restrict__ active, const MsLevel* L,
                                                 int32_t* touched, MsCounters* c, EdgeChunk* chunks) {
  if (L->done || L->bottom_up) return;
  const int64_t na = L->na;
  // A small frontier gets every vertex with >= kExpandBS edges chunked
  // across the grid; a large one only the real hubs.
  const int64_t chunk_min = na <= 4096 ? kExpandBS : kChunkEdges;
  __shared__ ExpandShared sh;
  __shared__ LdsQueue<kQueueCap> lq;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lq);
  TdVisit visit{dst, frontier, seen, acc, &lq, touched, &c->touched_tail};
  unsigned long long ex = 0;
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < na; tile += int64_t(gridDim.x) * kBS) {
    const int64_t i = tile + threadIdx.x;
    int32_t v = 0;
    int64_t beg = 0, end = 0;
    if (i < na) {
      v = active[i];
      beg = off[v];
      end = off[v + 1];
      ex += static_cast<unsigned long long>(end - beg);
    }
    ExpandTile(sh, v, 0, beg, end, visit, chunks, &c->nchunks, chunk_min);
  }
  LdsQueueFlush(lq, touched, &c->touched_tail);
  ex = BlockSum<kBS>(ex, s_red);
  if (threadIdx.x == 0 && ex) atomicAdd(&c->examined[blockIdx.x & (kStripes - 1)], ex);
}

__global__ __launch_bounds__(kBS) void MsTopDownChunks(const int32_t* __restrict__ dst,
                                                       const uint64_t* __restrict__ frontier,
                                                       const uint64_t* __restrict__ seen, uint64_t* acc,
                                                       int32_t* touched, const MsLevel* L, MsCounters* c,
                                                       const EdgeChunk* chunks) {
  if (L->done || L->bottom_up) return;
  __shared__ LdsQueue<kQueueCap> lq;
  LdsQueueInit(lq);
  TdVisit visit{dst, frontier, seen, acc, &lq, touched, &c->touched_tail};
  ExpandChunks(chunks, c->nchunks, visit);
  LdsQueueFlush(lq, touched, &c->touched_tail);
}

// Turns the accumulators of touched vertices into new bits and lists the
// vertices that gained any.
__global__ __launch_bounds__(kBS) void MsUpdate(const int32_t* __restrict__ touched, const MsCounters* cc,
                                                uint64_t* __restrict__ acc, uint64_t* __restrict__ seen,
                                                uint64_t* __restrict__ fnext, uint16_t* depth, const MsLevel* L,
                                                const int64_t* __restrict__ off,
                                                const int64_t* __restrict__ in_off, uint64_t full, int32_t* active,
                                                MsCounters* c) {
  if (L->done || L->bottom_up) return;
  const uint16_t level = static_cast<uint16_t>(L->level);
  __shared__ LdsQueue<kQueueCap> lq;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lq);
  const int32_t nt = cc->touched_tail;
  unsigned long long scout = 0, fin = 0, lv = 0, db = 0;
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < nt; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    bool take = false;
    int32_t w = 0;
    uint64_t s = 0, nb = 0;
    if (i < nt) {
      w = touched[i];
      s = seen[w];
      nb = acc[w] & ~s;
      acc[w] = 0;
      if (nb) {
        seen[w] = s | nb;
        fnext[w] = nb;
        scout += static_cast<unsigned long long>(off[w + 1] - off[w]);
        if (((s | nb) & full) == full) fin += static_cast<unsigned long long>(in_off[w + 1] - in_off[w]);
        lv |= nb;
        take = true;
      }
    }
    if (depth) {
      WaveWriteDepths(depth, w, nb, s, level);
      if (nb) db += DepthBytes(nb, s);
    }
    LdsQueuePush(lq, take, w, active, &c->active_tail);
  }
  LdsQueueFlush(lq, active, &c->active_tail);
  scout = BlockSum<kBS>(scout, s_red);
  fin = BlockSum<kBS>(fin, s_red);
  db = BlockSum<kBS>(db, s_red);
  lv = BlockOr<kBS>(lv, s_red);
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x & (kStripes - 1);
    if (db) atomicAdd(&c->depth_bytes[1][slot], db);
    if (scout) atomicAdd(&c->scout[slot], scout);
    if (fin) atomicAdd(&c->finished[slot], fin);
    if (lv) atomicOr(&c->live[slot], lv);
  }
}

// ---- bottom-up --------------------------------------------------------------
//
// `live` holds the roots with bits in the current frontier: a vertex can
// only gain bits of live roots, so it is skipped when it has all of them,
// and the heavy tier stops scanning once it has found them.

// Light tier: vertices below kHeavyEdges in-edges, 256 per block.  In each
// round, the next list positions of the tile's vertices that still need
// live bits are laid out as one flat range (block prefix sum); each thread
// gathers kUnroll neighbours' frontier masks per step, locating their owner
// by binary search over the prefix, and ORs them into the owner's LDS slot.
// A vertex leaves after the round that gave it every needed bit.  Writes
// fnext for every vertex without a pending live bit and every light one.
__global__ __launch_bounds__(kBS) void MsBottomUp(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                                  const int64_t* __restrict__ out_off,
                                                  const uint64_t* __restrict__ frontier, uint64_t* __restrict__ seen,
                                                  uint64_t* __restrict__ fnext, uint16_t* depth, int64_t n,
                                                  const MsLevel* L, uint64_t full, MsCounters* c,
                                                  uint64_t* __restrict__ acc) {
  if (L->done || !L->bottom_up) return;
  // After a propagation-block pass a light vertex's needed bits are in
  // acc[i] (taken here) and its list is not walked.
  const bool blocks = L->blocks;
  const uint16_t level = static_cast<uint16_t>(L->level);
  const uint64_t live = L->live;
  __shared__ int64_t s_beg[kBS];
  __shared__ int s_pre[kBS + 1];
  __shared__ unsigned long long s_got[kBS];
  __shared__ int s_scan[kBS / kWave];
  __shared__ unsigned long long s_red[kBS / kWave];
  const int tid = threadIdx.x;
  unsigned long long awake = 0, scout = 0, ex = 0, fin = 0, lv = 0, db = 0;
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < n; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + tid;
    uint64_t s = 0;
    int64_t b = 0;
    int d = 0, deg_full = 0;
    bool light = false;
    if (i < n) {
      s = seen[i];
      if ((live & ~s) == 0) {
        fnext[i] = 0;
      } else {
        b = in_off[i];
        const int64_t deg = in_off[i + 1] - b;
        if (deg < kHeavyEdges) {
          light = true;
          d = deg_full = static_cast<int>(deg);
        }
      }
    }
    s_got[tid] = 0;
    if (blocks && light) {
      d = 0;
      const uint64_t a = acc[i];
      if (a) {
        s_got[tid] = a;
        acc[i] = 0;
      }
    }
    const uint64_t need = live & ~s;
    // Rounds over list positions [kRoundEnd[r-1], kRoundEnd[r]): a vertex
    // takes part in a round only while it still lacks a needed bit (lists
    // are sorted, so in a degree-ordered layout the hubs come first).
    for (int round = 0, r_lo = 0; round < kRounds; r_lo = kRoundEnd[round], round++) {
      int take = 0;
      if (light && r_lo < d && (s_got[tid] & need) != need)
        take = (d < kRoundEnd[round] ? d : kRoundEnd[round]) - r_lo;
      s_beg[tid] = b + r_lo;
      int total = 0;
      const int rank = BlockExclusiveScan<kBS>(take, &total, s_scan);
      if (total == 0) break;  // block-uniform: no vertex has entries left
      s_pre[tid] = rank;
      if (tid == 0) s_pre[kBS] = total;
      __syncthreads();
      for (int j0 = 0; j0 < total; j0 += kBS * kUnroll) {
        int32_t nbr[kUnroll];
        int own[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const int j = j0 + u * kBS + tid;
          own[u] = -1;
          nbr[u] = 0;
          if (j < total) {
            int lo = 0, hi = kBS;  // s_pre[lo] <= j < s_pre[hi]
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (s_pre[mid] <= j)
                lo = mid;
              else
                hi = mid;
            }
            own[u] = lo;
            nbr[u] = in_dst[s_beg[lo] + (j - s_pre[lo])];
          }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          if (own[u] >= 0) {
            const uint64_t f = frontier[nbr[u]];
            if (f) atomicOr(&s_got[own[u]], static_cast<unsigned long long>(f));
          }
        }
      }
      if (tid == 0) ex += static_cast<unsigned long long>(total);
      __syncthreads();
    }
    uint64_t nb = 0;
    if (light) {
      nb = s_got[tid] & ~s;
      fnext[i] = nb;
      if (nb) {
        seen[i] = s | nb;
        scout += static_cast<unsigned long long>(out_off[i + 1] - out_off[i]);
        if (((s | nb) & full) == full) fin += static_cast<unsigned long long>(deg_full);
        lv |= nb;
        awake++;
      }
    }
    if (depth) {
      WaveWriteDepths(depth, static_cast<int32_t>(i), nb, s, level);
      if (nb) db += DepthBytes(nb, s);
    }
    __syncthreads();
  }
  awake = BlockSum<kBS>(awake, s_red);
  scout = BlockSum<kBS>(scout, s_red);
  ex = BlockSum<kBS>(ex, s_red);
  fin = BlockSum<kBS>(fin, s_red);
  db = BlockSum<kBS>(db, s_red);
  lv = BlockOr<kBS>(lv, s_red);
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x & (kStripes - 1);
    if (db) atomicAdd(&c->depth_bytes[0][slot], db);
    if (awake) atomicAdd(&c->awake[slot], awake);
    if (scout) atomicAdd(&c->scout[slot], scout);
    if (ex) atomicAdd(&c->examined[slot], ex);
    if (fin) atomicAdd(&c->finished[slot], fin);
    if (lv) atomicOr(&c->live[slot], lv);
  }
}

// Splits every heavy vertex's in-edges into chunks of kHeavyChunk (once per
// graph).
__global__ __launch_bounds__(kBS) void MsHeavyChunkBuild(const int64_t* __restrict__ in_off,
                                                         const int32_t* __restrict__ heavy, int32_t nheavy,
                                                         EdgeChunk* out, MsCounters* c) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < nheavy; i += int64_t(gridDim.x) * kBS) {
    const int32_t v = heavy[i];
    const int64_t b = in_off[v], e = in_off[v + 1];
    const int nc = static_cast<int>((e - b + kHeavyChunk - 1) / kHeavyChunk);
    const int base = atomicAdd(&c->heavy_chunks, nc);
    for (int k = 0; k < nc; k++) {
      const int64_t cb = b + k * kHeavyChunk;
      out[base + k] = EdgeChunk{cb, cb + kHeavyChunk < e ? cb + kHeavyChunk : e, v, 0};
    }
  }
}

// The first in-neighbour of every heavy chunk (the key the chunks are
// ordered by, once per graph).
__global__ __launch_bounds__(kBS) void MsChunkFirst(const int32_t* __restrict__ in_dst,
                                                    const EdgeChunk* __restrict__ chunks, int32_t nch,
                                                    int32_t* __restrict__ first) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < nch; i += int64_t(gridDim.x) * kBS)
    first[i] = in_dst[chunks[i].beg];
}

// Heavy tier: one wavefront per chunk of a vertex that still needs live
// bits; kUnroll x 64 in-neighbours per step, OR-reduced across the
// wavefront, stopping once this chunk and the others (acc, read without
// atomics: it only gains bits) hold every needed bit.  Results are ORed into
// acc[v] for MsHeavyFinish.
__global__ __launch_bounds__(kBS) void MsHeavyChunks(const int32_t* __restrict__ in_dst,
                                                     const uint64_t* __restrict__ frontier,
                                                     const uint64_t* __restrict__ seen, uint64_t* acc, const MsLevel* L,
                                                     const EdgeChunk* __restrict__ chunks, int32_t nch,
                                                     MsCounters* c) {
  // After a propagation-block pass acc already holds every needed bit.
  if (L->done || !L->bottom_up || L->blocks) return;
  const uint64_t live = L->live;
  __shared__ unsigned long long s_red[kBS / kWave];
  const int lane = LaneId();
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  unsigned long long ex = 0;
  for (int64_t it = wave; it < nch; it += nwaves) {
    const EdgeChunk ch = chunks[it];
    const uint64_t need = live & ~seen[ch.u];
    if (need == 0) continue;
    // Earlier chunks of the vertex may have found every needed bit already.
    uint64_t before = lane == 0 ? acc[ch.u] : 0;
    before = __shfl(before, 0);
    if ((before & need) == need) continue;
    const int64_t beg = ch.beg;
    uint64_t got = 0;
    for (int64_t base = beg; base < ch.end; base += kWave * kUnroll) {
      int32_t nbr[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; u++) {
        const int64_t k = base + u * kWave + lane;
        nbr[u] = k < ch.end ? in_dst[k] : -1;
      }
      uint64_t f = 0;
#pragma unroll
      for (int u = 0; u < kUnroll; u++)
        if (nbr[u] >= 0) f |= frontier[nbr[u]];
      got |= WaveOr(f);
      const int64_t stop = base + kWave * kUnroll;
      ex += static_cast<unsigned long long>((stop < ch.end ? stop : ch.end) - base);
      // The exit test must be uniform across the wavefront (WaveOr above
      // needs every lane): lane 0 reads acc and broadcasts it.
      uint64_t other = lane == 0 ? acc[ch.u] : 0;
      other = __shfl(other, 0);
      if (((got | other) & need) == need) break;
    }
    if (lane == 0 && (got & need)) atomicOr(reinterpret_cast<unsigned long long*>(&acc[ch.u]), got & need);
  }
  ex = BlockSum<kBS>(lane == 0 ? ex : 0ull, s_red);
  if (threadIdx.x == 0 && ex) atomicAdd(&c->heavy_ex[blockIdx.x & (kStripes - 1)], ex);
}

// Heavy tier, second pass: acc -> new bits, one thread per heavy vertex.
// Writes fnext for every heavy vertex.
__global__ __launch_bounds__(kBS) void MsHeavyFinish(const int32_t* __restrict__ heavy, int32_t nheavy,
                                                     uint64_t* __restrict__ seen, uint64_t* __restrict__ acc,
                                                     uint64_t* __restrict__ fnext, uint16_t* depth, const MsLevel* L,
                                                     const int64_t* __restrict__ out_off,
                                                     const int64_t* __restrict__ in_off, uint64_t full,
                                                     MsCounters* c) {
  if (L->done || !L->bottom_up) return;
  const uint16_t level = static_cast<uint16_t>(L->level);
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long awake = 0, scout = 0, fin = 0, lv = 0, db = 0;
  // Block-uniform trip count: the depth writer needs whole wavefronts.
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < nheavy; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    int32_t v = 0;
    uint64_t s = 0, nb = 0;
    if (i < nheavy) {
      v = heavy[i];
      const uint64_t a = acc[v];
      s = seen[v];
      nb = a & ~s;
      if (a) acc[v] = 0;
      fnext[v] = nb;
      if (nb) {
        seen[v] = s | nb;
        scout += static_cast<unsigned long long>(out_off[v + 1] - out_off[v]);
        if (((s | nb) & full) == full) fin += static_cast<unsigned long long>(in_off[v + 1] - in_off[v]);
        lv |= nb;
        awake++;
      }
    }
    if (depth) {
      WaveWriteDepths(depth, v, nb, s, level);
      if (nb) db += DepthBytes(nb, s);
    }
  }
  awake = BlockSum<kBS>(awake, s_red);
  scout = BlockSum<kBS>(scout, s_red);
  fin = BlockSum<kBS>(fin, s_red);
  db = BlockSum<kBS>(db, s_red);
  lv = BlockOr<kBS>(lv, s_red);
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x & (kStripes - 1);
    if (db) atomicAdd(&c->depth_bytes[2][slot], db);
    if (awake) atomicAdd(&c->awake[slot], awake);
    if (scout) atomicAdd(&c->scout[slot], scout);
    if (fin) atomicAdd(&c->finished[slot], fin);
    if (lv) atomicOr(&c->live[slot], lv);
  }
}

// Propagation-block pass of a dense bottom-up level.  MsStep sets L->blocks
// on the first bottom-up level after a top-down one while more than a set
// share of the in-edges belongs to vertices some root has not reached: that
// level reads almost every in-entry (no vertex has every bit yet, so the
// early exit rarely stops a list), and gathering 8-byte frontier masks at
// random moves a cache line for each.  Here the level runs through the
// layout Jacobi PageRank uses, with the masks as the values: every in-entry
// is read once, in two streams.
//   write  a block per chunk of a source tile stages the tile's masks in
//          LDS and writes each in-entry's source mask in block order
//          (16 bytes per lane);
//   or     a block per chunk of a destination bin ORs its entries' masks
//          into one LDS word per destination (a lane ORs its octet's runs
//          of equal destinations first), then stores, for each vertex that
//          still needs live bits, the needed bits it found in acc (or the
//          chunk's words, for a bin of several chunks);
//   merge  ORs those chunks' words per vertex into acc.
// The light tier then takes acc instead of walking its lists and the heavy
// tier stands down; MsHeavyFinish and MsBottomUp finish the level as usual.
constexpr int kMsBlkBS = 1024;

__global__ __launch_bounds__(kMsBlkBS) void MsBlocksWrite(const uint64_t* __restrict__ frontier, int64_t n, int lg,
                                                        const uint32_t* __restrict__ src16,
                                                        const PbChunk* __restrict__ chunks,
                                                        ulonglong2* __restrict__ val, const MsLevel* L) {
  if (L->done || !L->bottom_up || !L->blocks) return;
  __shared__ uint64_t slice[1 << 14];
  const PbChunk c = chunks[blockIdx.x];
  const int64_t base = int64_t(c.part) << lg;
  const int64_t cnt = min(int64_t(1) << lg, n - base);
  for (int64_t i = threadIdx.x; i < cnt; i += kMsBlkBS) slice[i] = frontier[base + i];
  __syncthreads();
  const uint32_t m = (1u << lg) - 1;  // keeps every slice index below 2^lg
  // Pairs of entries: one 32-bit word of offsets in, 16 bytes of masks out.
  const int64_t i1 = int64_t(c.q1) * 4;
  for (int64_t i = int64_t(c.q0) * 4 + threadIdx.x; i < i1; i += 2 * kMsBlkBS) {
    const bool two = i + kMsBlkBS < i1;
    const uint32_t w0 = src16[i], w1 = two ? src16[i + kMsBlkBS] : 0u;
    val[i] = make_ulonglong2(slice[w0 & m], slice[(w0 >> 16) & m]);
    if (two) val[i + kMsBlkBS] = make_ulonglong2(slice[w1 & m], slice[(w1 >> 16) & m]);
  }
}

__global__ __launch_bounds__(kMsBlkBS) void MsBlocksOr(const uint4* __restrict__ dst16, const uint32_t* __restrict__ oct,
                                                     const ulonglong2* __restrict__ val,
                                                     const PbChunk* __restrict__ chunks, int64_t n, int lg,
                                                     const uint64_t* __restrict__ seen, uint64_t* __restrict__ acc,
                                                     uint64_t* __restrict__ part, const MsLevel* L, MsCounters* c) {
  if (L->done || !L->bottom_up || !L->blocks) return;
  __shared__ unsigned long long s_acc[1 << 14];
  __shared__ unsigned long long s_red[kMsBlkBS / kWave];
  const PbChunk ch = chunks[blockIdx.x];
  const int w = 1 << lg;
  const uint32_t m = static_cast<uint32_t>(w - 1);
  for (int i = threadIdx.x; i < w; i += kMsBlkBS) s_acc[i] = 0;
  __syncthreads();
  unsigned long long ex = 0;
  for (int64_t q = int64_t(ch.q0) + threadIdx.x; q < ch.q1; q += kMsBlkBS) {
    const uint32_t o = oct[q];
    const uint4 dw = dst16[o];
    const uint32_t d[8] = {dw.x & 0xFFFFu, dw.x >> 16, dw.y & 0xFFFFu, dw.y >> 16,
                           dw.z & 0xFFFFu, dw.z >> 16, dw.w & 0xFFFFu, dw.w >> 16};
    const ulonglong2* p = val + 4 * int64_t(o);
    const ulonglong2 a = p[0], b = p[1], e = p[2], f = p[3];
    const uint64_t v[8] = {a.x, a.y, b.x, b.y, e.x, e.y, f.x, f.y};
    uint64_t run = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (d[i] == kPbPad) break;  // padding only follows the real entries
      ex++;
      run = (i > 0 && d[i] != d[i - 1] ? 0ull : run) | v[i];
      if ((i == 7 || d[i + 1] != d[i]) && run) atomicOr(&s_acc[d[i] & m], static_cast<unsigned long long>(run));
    }
  }
  __syncthreads();
  const int64_t first = int64_t(ch.part) << lg;
  const int cnt = static_cast<int>(min(int64_t(w), n - first));
  if (ch.slot < 0) {
    for (int i = threadIdx.x; i < cnt; i += kMsBlkBS) {
      const uint64_t got = s_acc[i];
      if (!got) continue;
      const uint64_t need = got & ~seen[first + i];  // frontier masks hold only live bits
      if (need) acc[first + i] = need;
    }
  } else {
    uint64_t* out = part + (int64_t(ch.slot) << lg);
    for (int i = threadIdx.x; i < cnt; i += kMsBlkBS) out[i] = s_acc[i];
  }
  ex = BlockSum<kMsBlkBS>(ex, s_red);
  if (threadIdx.x == 0 && ex) atomicAdd(&c->blk_ex[blockIdx.x & (kStripes - 1)], ex);
}

__global__ __launch_bounds__(kBS) void MsBlocksMerge(const PbMulti* __restrict__ multi, int32_t n_multi, int lg,
                                                     int64_t n, const uint64_t* __restrict__ part,
                                                     const uint64_t* __restrict__ seen, uint64_t* __restrict__ acc,
                                                     const MsLevel* L) {
  if (L->done || !L->bottom_up || !L->blocks) return;
  const int64_t w = int64_t(1) << lg, total = int64_t(n_multi) << lg;
  for (int64_t j = blockIdx.x * int64_t(kBS) + threadIdx.x; j < total; j += int64_t(gridDim.x) * kBS) {
    const PbMulti M = multi[j >> lg];
    const int64_t i = j & (w - 1), u = (int64_t(M.bin) << lg) + i;
    if (u >= n) continue;
    uint64_t got = 0;
    for (uint32_t s = 0; s < M.nslots; s++) got |= part[(int64_t(M.slot0 + s) << lg) + i];
    if (!got) continue;
    const uint64_t need = got & ~seen[u];
    if (need) acc[u] = need;
  }
}

// Dense frontier -> list, before a top-down level that follows a bottom-up
// one.
__global__ __launch_bounds__(kBS) void MsCompact(const uint64_t* __restrict__ frontier, int64_t n, int32_t* out,
                                                 const MsLevel* L, MsCounters* c) {
  if (L->done || L->bottom_up || L->listed) return;
  __shared__ LdsQueue<kQueueCap> lq;
  LdsQueueInit(lq);
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < n; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    LdsQueuePush(lq, i < n && frontier[i] != 0, static_cast<int32_t>(i), out, &c->compact_tail);
  }
  LdsQueueFlush(lq, out, &c->compact_tail);
}

// The clears MsStep asked for (L->clear) in the array the last level
// consumed; with L null, MsClearAll clears unconditionally.
__global__ __launch_bounds__(kBS) void MsClearList(uint64_t* __restrict__ frontier,
                                                   const int32_t* __restrict__ active, const MsLevel* L) {
  if (L->clear != 1) return;
  const int64_t na = L->na_prev;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < na; i += int64_t(gridDim.x) * kBS)
    frontier[active[i]] = 0;
}

__global__ __launch_bounds__(kBS) void MsClearAll(uint64_t* __restrict__ frontier, int64_t n, const MsLevel* L) {
  if (L && L->clear != 2) return;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    frontier[i] = 0;
}

// out[0] = sum over vertices of popcount(seen) (reached (root, vertex)
// pairs), out[1] = sum of out-degree x popcount(seen).
__global__ __launch_bounds__(kBS) void MsReach(const uint64_t* __restrict__ seen, const int64_t* __restrict__ off,
                                               int64_t n, unsigned long long* out) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long cnt = 0, deg = 0;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    const unsigned long long p = __popcll(seen[i]);
    cnt += p;
    deg += p * static_cast<unsigned long long>(off[i + 1] - off[i]);
  }
  cnt = BlockSum<kBS>(cnt, s_red);
  deg = BlockSum<kBS>(deg, s_red);
  if (threadIdx.x == 0) {
    if (cnt) atomicAdd(&out[0], cnt);
    if (deg) atomicAdd(&out[1], deg);
  }
}

__global__ void MsZero2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}


struct MsTraversal {
  uint64_t* seen = nullptr;          // final visit masks
  std::vector<uint64_t*> frontiers;  // keep mode: frontiers[L] holds the bits that arrived at level L
  int64_t td = 0, bu = 0, examined = 0, launches = 0;
  bool graph = false;                // levels replayed from a captured graph
};

// Device buffers and fixed arguments of one traversal.
struct MsBuffers {
  MsCounters* dc;
  MsLevel* L;
  const int32_t* heavy;
  const EdgeChunk* hchunks;
  int32_t nheavy, nhchunks;
  uint64_t* seen;
  uint64_t* acc;
  int32_t* touched;
  EdgeChunk* chunks;  // null: no chunking (no vertex reaches kExpandBS edges)
  uint16_t* depth16;
  uint64_t full;
  // Propagation blocks for the densest bottom-up level (pb null when the
  // graph runs without them): the layout, its arrays, the 8-byte value per
  // padded entry and the partial results of bins of several chunks.
  const gg_graph::PrBlocks* pb;
  PbArrays pba;
  uint64_t* pb_val;
  uint64_t* pb_part;
};

struct MsLaunched {
  int td = -1, update = -1, bu = -1, heavy = -1, heavy_finish = -1, blk_write = -1, blk_or = -1;  // profile records to credit once the counts are known
  int kernels = 0;
};

// Enqueues one level, fcur -> fnext (act lists fcur when it is listed;
// act_next receives a top-down level's new list): the top-down part, the
// bottom-up part and the compaction before a top-down level, each only
// when asked for.  Every kernel also checks the level state, so a captured
// graph can hold all of them.
inline MsLaunched MsLaunchLevel(gg_graph* g, const MsBuffers& b, const uint64_t* fcur, uint64_t* fnext,
                                int32_t* act, int32_t* act_next, bool td, bool bu, bool compact) {
  MsLaunched r;
  const int64_t n = g->n;
  const dim3 blk(kBS), grid_n(GridFor(n, kBS, g->cu_count * kLevelBlocksPerCu));
  const double N = static_cast<double>(n);
  if (compact) {
    Run(g, "msbfs_compact", 8 * N, MsCompact, dim3(GridFor(n, kBS, g->cu_count * 4)), blk, fcur, n, act,
        static_cast<const MsLevel*>(b.L), b.dc);
    r.kernels++;
  }
  if (td) {
    // 4 B per edge read, the neighbour's 8 B seen mask and 8 B accumulator.
    r.td = Run(g, "msbfs_top_down", 0, MsTopDown, dim3(g->cu_count * kLevelBlocksPerCu / 2), blk,
               static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst), fcur,
               static_cast<const uint64_t*>(b.seen), b.acc, static_cast<const int32_t*>(act),
               static_cast<const MsLevel*>(b.L), b.touched, b.dc, b.chunks);
    r.kernels++;
    if (b.chunks) {
      Run(g, "msbfs_top_down_chunks", 0, MsTopDownChunks, dim3(g->cu_count * kLevelBlocksPerCu / 4), blk,
          static_cast<const int32_t*>(g->out_dst), fcur, static_cast<const uint64_t*>(b.seen), b.acc, b.touched,
          static_cast<const MsLevel*>(b.L), b.dc, static_cast<const EdgeChunk*>(b.chunks));
      r.kernels++;
    }
    r.update = Run(g, "msbfs_update", 0, MsUpdate, dim3(g->cu_count * 4), blk, static_cast<const int32_t*>(b.touched),
        static_cast<const MsCounters*>(b.dc), b.acc, b.seen, fnext, b.depth16, static_cast<const MsLevel*>(b.L),
        static_cast<const int64_t*>(g->out_off), static_cast<const int64_t*>(g->in_off), b.full, act_next, b.dc);
    r.kernels++;
  }
  if (bu) {
    // Per in-edge gathered: 4 B and the neighbour's 8 B frontier mask
    // (credited by the caller).
    if (b.pb) {
      const gg_graph::PrBlocks& P = *b.pb;
      if (P.n_write) {
        r.blk_write = Run(g, "msbfs_blocks_write", 0, MsBlocksWrite, dim3(static_cast<unsigned>(P.n_write)), dim3(kMsBlkBS),
                          fcur, n, P.lg, reinterpret_cast<const uint32_t*>(b.pba.src16), b.pba.wchunks,
                          reinterpret_cast<ulonglong2*>(b.pb_val), static_cast<const MsLevel*>(b.L));
        r.kernels++;
      }
      r.blk_or = Run(g, "msbfs_blocks_or", 0, MsBlocksOr, dim3(static_cast<unsigned>(P.n_sum)), dim3(kMsBlkBS), b.pba.dst16,
                     b.pba.oct, reinterpret_cast<const ulonglong2*>(b.pb_val), b.pba.schunks, n, P.lg,
                     static_cast<const uint64_t*>(b.seen), b.acc, b.pb_part, static_cast<const MsLevel*>(b.L), b.dc);
      r.kernels++;
      if (P.n_multi) {
        Run(g, "msbfs_blocks_merge", 0, MsBlocksMerge, dim3(GridFor(P.n_multi << P.lg, kBS, g->cu_count * 8)), blk,
            b.pba.multi, static_cast<int32_t>(P.n_multi), P.lg, n, static_cast<const uint64_t*>(b.pb_part),
            static_cast<const uint64_t*>(b.seen), b.acc, static_cast<const MsLevel*>(b.L));
        r.kernels++;
      }
    }
    // Per vertex: seen (8 B) and fnext (8 B).
    r.bu = Run(g, "msbfs_bottom_up", 16 * N, MsBottomUp, grid_n, blk, static_cast<const int64_t*>(g->in_off),
               static_cast<const int32_t*>(g->in_dst), static_cast<const int64_t*>(g->out_off), fcur, b.seen,
               fnext, b.depth16, n, static_cast<const MsLevel*>(b.L), b.full, b.dc, b.acc);
    r.heavy = Run(g, "msbfs_bottom_up_heavy", 0, MsHeavyChunks,
                  dim3(GridFor(int64_t(b.nhchunks) * kWave, kBS, g->cu_count * kLevelBlocksPerCu)), blk,
                  static_cast<const int32_t*>(g->in_dst), fcur, static_cast<const uint64_t*>(b.seen), b.acc,
                  static_cast<const MsLevel*>(b.L), b.hchunks, b.nhchunks, b.dc);
    r.heavy_finish = Run(g, "msbfs_heavy_finish", 32.0 * b.nheavy, MsHeavyFinish, dim3(GridFor(b.nheavy, kBS, g->cu_count * 4)), blk,
        b.heavy, b.nheavy, b.seen, b.acc, fnext, b.depth16, static_cast<const MsLevel*>(b.L),
        static_cast<const int64_t*>(g->out_off), static_cast<const int64_t*>(g->in_off), b.full, b.dc);
    r.kernels += 3;
  }
  return r;
}

// Runs the level loop on g->stream from the `count` roots in droots
// (device memory, already validated).  Root i takes bit dbits[i] (device
// memory; `full` is the OR of the bits used), or bit i when dbits is null
// (count <= 64).  Roots sharing a bit must lie in different components.
// depth16 (n x 64, may be null) receives the depths per bit.  With keep,
// every level's frontier array is kept for the caller (workspace arrays
// named ms.level.<L>, valid until the next call).  The caller has set the
// device and started timing.
//
// Two drivers share the kernels and the device level state: the host loop
// (keep mode, or per-kernel profiling on) enqueues one level, waits for
// MsStep's mail and enqueues the next; otherwise two levels (one per
// frontier-buffer parity) are captured once into a hipGraph, cached in the
// handle, and replayed until the state says done, with one host read of
// the state per replay.
inline MsTraversal MsTraverse(gg_graph* g, const int32_t* droots, const int32_t* dbits, int count, uint64_t full,
                              int alpha, uint16_t* depth16, bool keep) {
  MsTraversal res;
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const dim3 blk(kBS);
  MsCounters* dc = g->Scratch<MsCounters>("ms.ctr", 1);
  MsLevel* dl = g->Scratch<MsLevel>("ms.level_state", 1);
  int32_t* heavy = g->Scratch<int32_t>("ms.heavy", n + 1);
  // Heavy vertices and their in-edge chunks depend only on the graph.
  const int64_t max_chunks = g->in_entries / kHeavyChunk + g->in_entries / kHeavyEdges + 1;
  EdgeChunk* hchunks = g->Scratch<EdgeChunk>("ms.hchunks", max_chunks);
  if (g->ms_heavy < 0) {
    GG_HIP(hipMemsetAsync(dc, 0, sizeof(MsCounters), st));
    MsHeavyList<<<dim3(GridFor(n, kBS, g->cu_count * 4)), blk, 0, st>>>(g->in_off, n, heavy, dc);
    GG_HIP(hipGetLastError());
    int32_t nh = 0;
    GG_HIP(hipMemcpyAsync(&nh, &dc->heavy_tail, sizeof(nh), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    MsHeavyChunkBuild<<<dim3(GridFor(nh, kBS, g->cu_count * 4)), blk, 0, st>>>(g->in_off, heavy, nh, hchunks,
                                                                                dc);
    GG_HIP(hipGetLastError());
    int32_t nc = 0;
    GG_HIP(hipMemcpyAsync(&nc, &dc->heavy_chunks, sizeof(nc), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    GG_REQUIRE(nc <= max_chunks, GG_EHIP, "heavy chunk count exceeds its bound");
    // Chunks in the order of their first in-neighbour (ties by vertex).
    // Lists are sorted, so chunks running side by side start among the same
    // sources and share those frontier masks in L2, and every heavy
    // vertex's first chunk -- its hub neighbours, where it most likely finds
    // its bits -- runs before its later ones, which can then stop at once.
    if (nc > 1) {
      int32_t* first = g->Scratch<int32_t>("ms.hchunk_first", nc);
      MsChunkFirst<<<dim3(GridFor(nc, kBS, g->cu_count * 4)), blk, 0, st>>>(g->in_dst, hchunks, nc, first);
      GG_HIP(hipGetLastError());
      std::vector<EdgeChunk> hc(nc);
      std::vector<int32_t> key(nc);
      GG_HIP(hipMemcpyAsync(hc.data(), hchunks, sizeof(EdgeChunk) * nc, hipMemcpyDeviceToHost, st));
      GG_HIP(hipMemcpyAsync(key.data(), first, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      std::vector<int32_t> idx(nc);
      for (int32_t i = 0; i < nc; i++) idx[i] = i;
      std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        return key[a] != key[b] ? key[a] < key[b] : (hc[a].u != hc[b].u ? hc[a].u < hc[b].u : hc[a].beg < hc[b].beg);
      });
      std::vector<EdgeChunk> sorted(nc);
      for (int32_t i = 0; i < nc; i++) sorted[i] = hc[idx[i]];
      GG_HIP(hipMemcpyAsync(hchunks, sorted.data(), sizeof(EdgeChunk) * nc, hipMemcpyHostToDevice, st));
      GG_HIP(hipStreamSynchronize(st));
      g->ws.Drop("ms.hchunk_first");
    }
    g->ms_heavy = nh;
    g->ms_heavy_chunks = nc;
  }
  uint64_t* seen = g->Scratch<uint64_t>("ms.seen", n);
  // keep: one frontier array per level, else two alternating ones.
  auto level_array = [&](int level) {
    if (!keep) return g->Scratch<uint64_t>((level & 1) ? "ms.fb" : "ms.fa", n);
    return g->Scratch<uint64_t>(("ms.level." + std::to_string(level)).c_str(), n);
  };
  uint64_t* fa = level_array(0);
  uint64_t* fb = level_array(1);
  if (keep) res.frontiers.push_back(fa);
  uint64_t* acc = g->Scratch<uint64_t>("ms.acc", n);
  int32_t* act_a = g->Scratch<int32_t>("ms.act_a", n + 64);
  int32_t* act_b = g->Scratch<int32_t>("ms.act_b", n + 64);
  int32_t* touched = g->Scratch<int32_t>("ms.touched", n + 64);
  const bool use_chunks = g->max_degree >= kExpandBS;
  EdgeChunk* chunks = use_chunks ? g->Scratch<EdgeChunk>("ms.chunks", MaxChunks(n, g->out_entries)) : nullptr;
  // Propagation blocks (the layout FP32 Jacobi PageRank runs through, built
  // on first use) for the first bottom-up level after a top-down one while
  // more than bfs_blocks_pct percent (default 50) of the in-edges are
  // unfinished.  The values share FP64 PageRank's arrays (8 bytes each).
  // By default only for batches of 32 roots or more: with fewer roots the
  // gathers' early exit stops lists sooner (Kronecker 22, 16 roots: 1.54 ms
  // without the pass, 2.09 ms with it).
  const bool blocks_wanted = g->tune.bfs_blocks_pct > 0 || (g->tune.bfs_blocks_pct == 0 && __builtin_popcountll(full) >= 32);
  const gg_graph::PrBlocks* PB = blocks_wanted ? PropagationBlocks(g) : nullptr;
  PbArrays pba{};
  uint64_t* pb_val = nullptr;
  uint64_t* pb_part = nullptr;
  int64_t blocks_min = INT64_MAX;
  if (PB) {
    pba = PropagationBlockArrays(g);
    pb_val = g->Scratch<uint64_t>("pr.pb.val64", 8 * PB->octets + 8);
    pb_part = g->Scratch<uint64_t>("pr.pb.part64", std::max<int64_t>(1, PB->slots << PB->lg));
    const int pct = g->tune.bfs_blocks_pct > 0 ? g->tune.bfs_blocks_pct : 50;
    blocks_min = pct == 100 ? g->in_entries - 1 : g->in_entries * pct / 100;
  }
  const MsBuffers b{dc,   dl,  heavy,   hchunks, static_cast<int32_t>(g->ms_heavy), static_cast<int32_t>(g->ms_heavy_chunks),
                    seen, acc, touched, chunks,  depth16,                            full,
                    PB, pba, pb_val, pb_part};
  const dim3 grid_n(GridFor(n, kBS, g->cu_count * 16));
  const double N = static_cast<double>(n);
  int64_t& launches = res.launches;
  unsigned long long mail[kMailValues];

  Run(g, "msbfs_init", 32 * N, MsInit, grid_n, blk, seen, fa, fb, acc, n, dc);
  Run(g, "msbfs_roots", 0, MsRoots, dim3(GridFor(count, 64)), dim3(64), droots, dbits, count, seen, fa, depth16,
      act_a, dc);
  Run(g, "msbfs_level_init", 0, MsLevelInit, dim3(1), dim3(64), dc, dl, g->in_entries, full, blocks_min);
  launches += 3;

  MsLevel hl;
  if (!keep && !g->profile) {
    // Graph-replayed levels.
    const std::vector<uintptr_t> key = {
        reinterpret_cast<uintptr_t>(fa),      reinterpret_cast<uintptr_t>(fb),      reinterpret_cast<uintptr_t>(act_a),
        reinterpret_cast<uintptr_t>(act_b),   reinterpret_cast<uintptr_t>(dc),      reinterpret_cast<uintptr_t>(dl),
        reinterpret_cast<uintptr_t>(seen),    reinterpret_cast<uintptr_t>(acc),     reinterpret_cast<uintptr_t>(touched),
        reinterpret_cast<uintptr_t>(chunks),  reinterpret_cast<uintptr_t>(hchunks), reinterpret_cast<uintptr_t>(heavy),
        reinterpret_cast<uintptr_t>(depth16), static_cast<uintptr_t>(full),         static_cast<uintptr_t>(alpha),
        static_cast<uintptr_t>(g->ms_heavy),  static_cast<uintptr_t>(g->ms_heavy_chunks),
        reinterpret_cast<uintptr_t>(PB), reinterpret_cast<uintptr_t>(pba.src16), reinterpret_cast<uintptr_t>(pba.dst16),
        reinterpret_cast<uintptr_t>(pba.oct), reinterpret_cast<uintptr_t>(pba.wchunks),
        reinterpret_cast<uintptr_t>(pba.schunks), reinterpret_cast<uintptr_t>(pba.multi),
        reinterpret_cast<uintptr_t>(pb_val), reinterpret_cast<uintptr_t>(pb_part),
        static_cast<uintptr_t>(PB ? PB->lg : 0), static_cast<uintptr_t>(PB ? PB->n_write : 0),
        static_cast<uintptr_t>(PB ? PB->n_sum : 0), static_cast<uintptr_t>(PB ? PB->n_multi : 0)};
    int per_replay = 0;
    if (!g->ms_exec || g->ms_exec_key != key) {
      if (g->ms_exec) {
        (void)hipGraphExecDestroy(g->ms_exec);
        g->ms_exec = nullptr;
      }
      hipGraph_t graph = nullptr;
      GG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        for (int half = 0; half < 2; half++) {
          uint64_t* fc = half ? fb : fa;
          uint64_t* fn = half ? fa : fb;
          int32_t* ac = half ? act_b : act_a;
          int32_t* an = half ? act_a : act_b;
          per_replay += MsLaunchLevel(g, b, fc, fn, ac, an, true, true, true).kernels;
          hipLaunchKernelGGL(MsStep, dim3(1), dim3(64), 0, st, dc, dl, alpha, 0, static_cast<unsigned long long*>(nullptr),
                             0ull);
          hipLaunchKernelGGL(MsClearList, dim3(g->cu_count * 8), blk, 0, st, fc, static_cast<const int32_t*>(ac),
                             static_cast<const MsLevel*>(dl));
          hipLaunchKernelGGL(MsClearAll, grid_n, blk, 0, st, fc, n, static_cast<const MsLevel*>(dl));
          per_replay += 3;
        }
        GG_HIP(hipGetLastError());
      } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      GG_HIP(hipStreamEndCapture(st, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      GG_HIP(e);
      g->ms_exec = exec;
      g->ms_exec_key = key;
      g->ms_exec_kernels = per_replay;
    }
    per_replay = g->ms_exec_kernels;
    res.graph = true;
    for (;;) {
      GG_HIP(hipGraphLaunch(g->ms_exec, st));
      launches += per_replay;
      GG_HIP(hipMemcpyAsync(&hl, dl, sizeof(hl), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      if (hl.done) break;
      GG_REQUIRE(hl.level < 65534, GG_ENOTSUP, "more than 65533 levels");
    }
  } else {
    // Host-driven levels.
    GG_HIP(hipMemcpyAsync(&hl, dl, sizeof(hl), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    int64_t na = hl.na;
    bool listed = true, bottom_up = false;
    uint64_t* fcur = fa;
    uint64_t* fnext = fb;
    int32_t* act = act_a;
    int32_t* act_next = act_b;
    for (int level = 1; na > 0; level++) {
      GG_REQUIRE(level < 65535, GG_ENOTSUP, "more than 65534 levels");
      if (keep) {
        // A fresh array per level; a top-down level writes only the entries
        // that gain bits, so it starts from zeros.
        fnext = level_array(level);
        res.frontiers.push_back(fnext);
        if (!bottom_up) {
          Run(g, "msbfs_clear", 8 * N, MsClearAll, grid_n, blk, fnext, n, static_cast<const MsLevel*>(nullptr));
          launches++;
        }
      }
      const MsLaunched r = MsLaunchLevel(g, b, fcur, fnext, act, act_next, !bottom_up, bottom_up,
                                         !bottom_up && !listed);
      launches += r.kernels;
      const unsigned long long seq = MailNext(g);
      Run(g, "msbfs_step", 0, MsStep, dim3(1), dim3(64), dc, dl, alpha, keep ? 1 : 0, g->mail_dev, seq);
      launches++;
      MailWait(g, seq, mail, kMailValues);
      if (r.td >= 0) CreditBytes(g, r.td, 20.0 * mail[kMailExamined]);
      if (r.bu >= 0) CreditBytes(g, r.bu, 12.0 * mail[kMailExamined] + mail[kMailDepthBottomUp]);
      if (r.update >= 0) CreditBytes(g, r.update, static_cast<double>(mail[kMailDepthUpdate]));
      if (r.heavy_finish >= 0) CreditBytes(g, r.heavy_finish, static_cast<double>(mail[kMailDepthHeavy]));
      if (r.heavy >= 0) CreditBytes(g, r.heavy, 12.0 * mail[kMailHeavyEx]);
      if (mail[kMailBlocksEx] && b.pb) {  // the blocks pass ran: its streams per padded entry, and 8 B per vertex
        const double ent = 8.0 * b.pb->octets;
        if (r.blk_write >= 0) CreditBytes(g, r.blk_write, 10.0 * ent + 8 * N);
        if (r.blk_or >= 0) CreditBytes(g, r.blk_or, 10.5 * ent + 8 * N);
      }
      const int64_t na_next = static_cast<int64_t>(mail[kMailNa]);
      const bool next_bu = mail[kMailBottomUp] != 0;
      // The clears MsStep chose for the array just consumed.
      if (!keep && na_next > 0 && !next_bu) {
        if (listed)
          Run(g, "msbfs_clear", 12.0 * na, MsClearList, dim3(GridFor(na, kBS, g->cu_count * 8)), blk, fcur,
              static_cast<const int32_t*>(act), static_cast<const MsLevel*>(dl));
        else
          Run(g, "msbfs_clear", 8 * N, MsClearAll, grid_n, blk, fcur, n, static_cast<const MsLevel*>(dl));
        launches++;
      }
      listed = !bottom_up;
      bottom_up = next_bu;
      na = na_next;
      std::swap(fcur, fnext);
      std::swap(act, act_next);
    }
    GG_HIP(hipMemcpyAsync(&hl, dl, sizeof(hl), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
  }
  res.td = hl.td;
  res.bu = hl.bu;
  res.examined = hl.examined;
  res.seen = seen;
  return res;
}

}  // namespace
}  // namespace gg
