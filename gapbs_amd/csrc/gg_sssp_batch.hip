// Batched delta-stepping: up to 64 single-source shortest path searches in
// one traversal on gfx950.
//
// Reference counterpart: the trial loop over roots (src/benchmark.h:106)
// around DeltaStep / RelaxEdges (src/sssp.cc:64-153), which runs one root per
// trial.  Kept from DeltaStep: int32 weights and distances, kDistInf =
// INT32_MAX / 2 for unreached vertices (sssp.cc:64), buckets of width delta
// processed in increasing order, relaxation as a compare-and-set minimum
// that re-queues an improved vertex (sssp.cc:68-85), and the current bucket
// processed again until it stays empty (the effect of the bucket fusion,
// sssp.cc:113-120).  Shortest distances are unique, so every root's result
// equals the reference's bit for bit whatever the processing order.
//
// Layout and work mapping.  Each vertex holds a row of G int32 distances
// (G = the smallest power of two >= k; entry r is the distance from root r),
// so one root's distances sit G * 4 bytes apart and a vertex's row is one
// or two cache lines.  A wavefront is cut into 64 / G groups of G lanes and
// a group works on one vertex at a time, lane r for root r: the edge list is
// read once per vertex for every root the vertex is active for, and the
// neighbour rows are read and updated by adjacent lanes.
//
// Pending work.  chg[v] (64 bits) holds the roots whose distance at v
// improved and has not been expanded yet.  An improvement lowers the
// distance with atomicMin, then sets the root's bit with atomicOr and queues
// v once per launch: on `near` if the new distance lies in the current
// bucket, else on the `fresh` pile.  Expanding u takes chg[u] with an atomic
// exchange, relaxes the edges for the taken roots whose distance lies in the
// current bucket and puts the others back.  Since a bit is set only after its
// distance is written and taken before the distance is read, no improvement
// is lost, whatever the interleaving.
//
// Buckets.  The buckets are shared by the roots: bucket [lo, hi) expands
// every (vertex, root) whose pending distance lies in it.  When near stays
// empty, the next bucket starts at the smallest pending distance, known
// from the minima the relax launches took over their fresh pushes and the
// split took over the pile it kept; a split then moves the pile's entries
// with a pending distance in the new bucket to the frontier and keeps those
// with a pending distance beyond it (each vertex once, by its stamps).
//
// Device-resident loop.  All of the loop state (phase, bucket, queue sizes,
// tags, minima) lives in device memory and two one-block step kernels
// advance it; every other kernel reads it and returns at once unless its
// phase runs.  A round is [relax, relax chunks, step, split, step]; rounds
// are captured into a hipGraph once per handle and replayed until the state
// says done, with one 128-byte state read per replay.
#include <climits>
#include <vector>

#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kQueueCap = 1024;
constexpr int32_t kDistInf = INT32_MAX / 2;
constexpr int kStripes = 64;
// Vertices with at least kSerialMax edges are cut into kChunkE-edge chunks
// that the chunk kernel hands to groups across the grid.
constexpr int64_t kSerialMax = 256;
constexpr int64_t kChunkE = 256;
constexpr int kUnroll = 4;  // edges whose loads a group keeps in flight
constexpr unsigned kBlocksPerCu = 8;
constexpr int kRoundsPerReplay = 8;

enum Phase { kRelax = 0, kSplit = 1, kDone = 2, kFailed = 3 };

struct SbChunk {
  int64_t beg, end;
  int32_t u, pad;
  uint64_t mask;  // the roots expanded
};

// Loop state.  Written by the step kernels only (and by SbInit); the other
// kernels append through the tails and stripes.
struct SbState {
  int32_t phase;
  int32_t parity;      // frontier = q[parity], near = q[parity ^ 1]
  int32_t pile_par;    // pile = p[pile_par], the split's kept pile = p[pile_par ^ 1]
  int32_t fsize;       // frontier entries
  int32_t pile_size;   // pile entries
  int32_t tag;         // last stamp tag handed out
  int32_t near_tag, fresh_tag, f_tag, pile_tag;
  int64_t lo, hi;      // current bucket
  int64_t delta;
  int64_t relax_launches, buckets, max_launches;
  int64_t relaxed, improved, scanned;
  int64_t frontier_in, split_in;  // frontier entries expanded, pile entries split
  // appended to by the kernels
  int32_t near_tail, fresh_tail, f_tail, pile2_tail, nchunks;
  int32_t pad;
  int32_t fresh_min[kStripes];  // smallest distance pushed to fresh since the last split
  int32_t pile_min[kStripes];   // smallest pending distance the last split kept
  unsigned long long sums[kStripes][3];  // relaxations, improvements, edges scanned
};

struct Group {
  int lg, r, leader;
  uint64_t gmask;
};

__device__ __forceinline__ Group MakeGroup(int lg) {
  Group g;
  const int lane = LaneId();
  g.lg = lg;
  g.r = lane & ((1 << lg) - 1);
  g.leader = lane & ~((1 << lg) - 1);
  const uint64_t ones = lg == 6 ? ~0ull : ((1ull << (1 << lg)) - 1);
  g.gmask = ones << g.leader;
  return g;
}

// This group's lanes with p, as root bits.  Call with the whole group active.
__device__ __forceinline__ uint64_t GroupBallot(const Group& g, bool p) {
  return (static_cast<uint64_t>(__ballot(p)) & g.gmask) >> g.leader;
}

__device__ __forceinline__ int32_t LoadC(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBS) void SbInitRows(int32_t* __restrict__ dist, uint64_t* __restrict__ chg,
                                                  int32_t* __restrict__ stamp, int32_t* __restrict__ freshstamp,
                                                  int32_t* __restrict__ pilestamp, int64_t n, int lg) {
  const int64_t rows = n << lg;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < rows; i += int64_t(gridDim.x) * kBS) {
    dist[i] = kDistInf;
    if (i < n) {
      chg[i] = 0;
      stamp[i] = -1;
      freshstamp[i] = -1;
      pilestamp[i] = -1;
    }
  }
}

// Roots: distance 0, pending bit, and the distinct root vertices as the
// first frontier (stamp tag 1).  One thread per root; roots may repeat.
__global__ void SbRoots(const int32_t* __restrict__ roots, int k, int lg, int32_t* __restrict__ dist,
                        uint64_t* __restrict__ chg, int32_t* __restrict__ stamp, int32_t* __restrict__ q0,
                        SbState* S) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const int32_t u = roots[r];
  dist[(int64_t(u) << lg) + r] = 0;
  atomicOr(reinterpret_cast<unsigned long long*>(&chg[u]), 1ull << r);
  if (atomicExch(&stamp[u], 1) != 1) q0[atomicAdd(&S->fsize, 1)] = u;
}

__global__ void SbStateInit(SbState* S, int64_t delta, int64_t max_launches) {
  const int t = threadIdx.x;
  if (t < kStripes) {
    S->fresh_min[t] = kDistInf;
    S->pile_min[t] = kDistInf;
    S->sums[t][0] = S->sums[t][1] = S->sums[t][2] = 0;
  }
  if (t == 0) {
    S->phase = kRelax;
    S->parity = 0;
    S->pile_par = 0;
    S->fsize = 0;
    S->pile_size = 0;
    // Tag 1 is the roots' frontier stamp (SbRoots); 2 and 3 the first
    // launch's near and fresh stamps.
    S->near_tag = 2;
    S->fresh_tag = 3;
    S->f_tag = S->pile_tag = 0;
    S->tag = 3;
    S->lo = 0;
    S->hi = delta;
    S->delta = delta;
    S->relax_launches = S->buckets = 0;
    S->max_launches = max_launches;
    S->relaxed = S->improved = S->scanned = 0;
    S->frontier_in = S->split_in = 0;
    S->near_tail = S->fresh_tail = S->f_tail = S->pile2_tail = S->nchunks = 0;
  }
}

struct RelaxCtx {
  const int2* wdst;
  int32_t* dist;
  uint64_t* chg;
  int32_t* stamp;
  int32_t* freshstamp;
  int32_t* near;
  int32_t* fresh;
  SbState* S;
  LdsQueue<kQueueCap>* lnear;
  LdsQueue<kQueueCap>* lfresh;
  int64_t hi;
  int32_t near_tag, fresh_tag;
  int lg;
  unsigned long long relaxed, improved, scanned;
  int32_t fmin;
};

// Relaxes edges [b, e) of u for the lanes with `active` (distance du).
// Called by whole groups; groups of one wavefront may diverge.
__device__ __forceinline__ void RelaxRange(RelaxCtx& c, const Group& g, int32_t u, int64_t b, int64_t e, bool active,
                                           int32_t du) {
  (void)u;
  for (int64_t base = b; base < e; base += kUnroll) {
    int2 vw[kUnroll];
    int32_t cur[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
      vw[j] = base + j < e ? c.wdst[base + j] : make_int2(0, 0);
    }
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
      cur[j] = active && base + j < e ? c.dist[(int64_t(vw[j].x) << c.lg) + g.r] : 0;
    }
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
      if (base + j >= e) break;  // group-uniform
      const int32_t v = vw[j].x;
      bool imp = false, to_near = false;
      int32_t nd32 = 0;
      if (active) {
        c.relaxed++;
        const int64_t nd = static_cast<int64_t>(du) + vw[j].y;
        if (nd < kDistInf && nd < cur[j]) {
          nd32 = static_cast<int32_t>(nd);
          const int32_t old = atomicMin(&c.dist[(int64_t(v) << c.lg) + g.r], nd32);
          imp = nd32 < old;
          to_near = imp && nd < c.hi;
          if (imp && !to_near && nd32 < c.fmin) c.fmin = nd32;
        }
      }
      const uint64_t bits = GroupBallot(g, imp);
      const uint64_t nbits = GroupBallot(g, to_near);
      bool push_near = false, push_fresh = false;
      if (bits && LaneId() == g.leader) {
        c.improved += __popcll(bits);
        atomicOr(reinterpret_cast<unsigned long long*>(&c.chg[v]), static_cast<unsigned long long>(bits));
        if (nbits) push_near = atomicExch(&c.stamp[v], c.near_tag) != c.near_tag;
        if (bits & ~nbits) push_fresh = atomicExch(&c.freshstamp[v], c.fresh_tag) != c.fresh_tag;
      }
      LdsQueuePush(*c.lnear, push_near, v, c.near, &c.S->near_tail);
      LdsQueuePush(*c.lfresh, push_fresh, v, c.fresh, &c.S->fresh_tail);
    }
    if (LaneId() == g.leader) c.scanned += (e - base < kUnroll ? e - base : kUnroll);
  }
}

__device__ __forceinline__ void RelaxFlush(RelaxCtx& c, unsigned long long* s_red, int32_t* s_min) {
  LdsQueueFlush(*c.lnear, c.near, &c.S->near_tail);
  LdsQueueFlush(*c.lfresh, c.fresh, &c.S->fresh_tail);
  const unsigned long long r = BlockSum<kBS>(c.relaxed, s_red);
  const unsigned long long im = BlockSum<kBS>(c.improved, s_red);
  const unsigned long long sc = BlockSum<kBS>(c.scanned, s_red);
  int32_t m = c.fmin;
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(m, o);
    m = t < m ? t : m;
  }
  if (LaneId() == 0) s_min[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBS / kWave; w++) m = s_min[w] < m ? s_min[w] : m;
    const int stripe = blockIdx.x & (kStripes - 1);
    if (m < kDistInf) atomicMin(&c.S->fresh_min[stripe], m);
    if (r) atomicAdd(&c.S->sums[stripe][0], r);
    if (im) atomicAdd(&c.S->sums[stripe][1], im);
    if (sc) atomicAdd(&c.S->sums[stripe][2], sc);
  }
}

// Expands frontier vertex u with its whole group: take its pending roots,
// put back those beyond the bucket, relax the rest (hubs: hand their edges
// to the chunk list).  Group-uniform control flow.
__device__ __forceinline__ void GroupExpand(RelaxCtx& c, const Group& g, int32_t u, const int64_t* __restrict__ off,
                                            SbChunk* __restrict__ chunks, SbState* S, int lg) {
  unsigned long long m = 0;
  if (LaneId() == g.leader) m = atomicExch(reinterpret_cast<unsigned long long*>(&c.chg[u]), 0ull);
  m = __shfl(m, g.leader);
  const bool mine = (m >> g.r) & 1;
  const int32_t du = mine ? LoadC(&c.dist[(int64_t(u) << lg) + g.r]) : kDistInf;
  const bool active = mine && du < c.hi;
  const uint64_t later = GroupBallot(g, mine && !active);
  const uint64_t act = GroupBallot(g, active);
  if (later && LaneId() == g.leader)
    atomicOr(reinterpret_cast<unsigned long long*>(&c.chg[u]), static_cast<unsigned long long>(later));
  if (!act) return;
  const int64_t b = off[u], e = off[u + 1];
  if (e - b >= kSerialMax) {
    if (LaneId() == g.leader) {
      const int64_t nc = (e - b + kChunkE - 1) / kChunkE;
      const int at = atomicAdd(&S->nchunks, static_cast<int>(nc));
      for (int64_t x = 0; x < nc; x++) {
        const int64_t cb = b + x * kChunkE;
        chunks[at + x] = SbChunk{cb, cb + kChunkE < e ? cb + kChunkE : e, u, 0, act};
      }
    }
    return;
  }
  RelaxRange(c, g, u, b, e, active, du);
}

// Expands frontier vertex u with one lane, for the few roots pending there
// (the same rules as GroupExpand, root by root).  Lanes of a wavefront may
// diverge; the queue pushes are wave-wide over the lanes still active.
__device__ __forceinline__ void LaneExpand(RelaxCtx& c, int32_t u, const int64_t* __restrict__ off,
                                           SbChunk* __restrict__ chunks, SbState* S, int lg) {
  const uint64_t m = atomicExch(reinterpret_cast<unsigned long long*>(&c.chg[u]), 0ull);
  const int64_t b = off[u], e = off[u + 1];
  const bool hub = e - b >= kSerialMax;
  uint64_t later = 0, act = 0;
  for (uint64_t rest = m; rest; rest &= rest - 1) {
    const int r = __builtin_ctzll(rest);
    const int32_t du = LoadC(&c.dist[(int64_t(u) << lg) + r]);
    if (du >= c.hi) {
      later |= 1ull << r;
      continue;
    }
    act |= 1ull << r;
    if (hub) continue;
    for (int64_t k = b; k < e; k++) {
      const int2 vw = c.wdst[k];
      const int32_t v = vw.x;
      const int64_t nd = static_cast<int64_t>(du) + vw.y;
      int32_t* dv = &c.dist[(int64_t(v) << lg) + r];
      bool push_near = false, push_fresh = false;
      c.relaxed++;
      if (nd < kDistInf && nd < *dv) {
        const int32_t nd32 = static_cast<int32_t>(nd);
        if (nd32 < atomicMin(dv, nd32)) {
          c.improved++;
          atomicOr(reinterpret_cast<unsigned long long*>(&c.chg[v]), 1ull << r);
          if (nd < c.hi) {
            push_near = atomicExch(&c.stamp[v], c.near_tag) != c.near_tag;
          } else {
            push_fresh = atomicExch(&c.freshstamp[v], c.fresh_tag) != c.fresh_tag;
            if (nd32 < c.fmin) c.fmin = nd32;
          }
        }
      }
      LdsQueuePush(*c.lnear, push_near, v, c.near, &c.S->near_tail);
      LdsQueuePush(*c.lfresh, push_fresh, v, c.fresh, &c.S->fresh_tail);
    }
    c.scanned += static_cast<unsigned long long>(e - b);
  }
  if (later) atomicOr(reinterpret_cast<unsigned long long*>(&c.chg[u]), static_cast<unsigned long long>(later));
  if (hub && act) {
    const int64_t nc = (e - b + kChunkE - 1) / kChunkE;
    const int at = atomicAdd(&S->nchunks, static_cast<int>(nc));
    for (int64_t x = 0; x < nc; x++) {
      const int64_t cb = b + x * kChunkE;
      chunks[at + x] = SbChunk{cb, cb + kChunkE < e ? cb + kChunkE : e, u, 0, act};
    }
  }
}

// Expands the frontier q[parity].  With groups of 8 or more lanes, a
// vertex with few pending roots (at most G / 16, and at least one) would
// leave most of its group idle -- on graphs whose roots rarely share a
// bucket at a vertex (grids, uniform graphs) that is nearly every vertex --
// so the block first reads its tile's pending masks: those vertices are
// expanded one lane each, the others queued in LDS for a group each.
__global__ __launch_bounds__(kBS) void SbRelax(const int64_t* __restrict__ off, const int2* __restrict__ wdst,
                                               int32_t* __restrict__ dist, uint64_t* __restrict__ chg,
                                               int32_t* __restrict__ stamp, int32_t* __restrict__ freshstamp,
                                               int32_t* __restrict__ qa, int32_t* __restrict__ qb,
                                               int32_t* __restrict__ fresh, SbChunk* __restrict__ chunks, SbState* S,
                                               int lg) {
  if (S->phase != kRelax) return;
  __shared__ LdsQueue<kQueueCap> lnear, lfresh;
  __shared__ unsigned long long s_red[kBS / kWave];
  __shared__ int32_t s_min[kBS / kWave];
  __shared__ int32_t s_dense[kBS];
  __shared__ int s_ndense;
  LdsQueueInit(lnear);
  LdsQueueInit(lfresh);
  const int parity = S->parity;
  const int32_t* F = parity ? qb : qa;
  int32_t* near = parity ? qa : qb;
  const int64_t fsize = S->fsize;
  RelaxCtx c{wdst, dist, chg, stamp, freshstamp, near, fresh, S, &lnear, &lfresh, S->hi, S->near_tag,
             S->fresh_tag, lg, 0, 0, 0, kDistInf};
  const Group g = MakeGroup(lg);
  const int groups = kWave >> lg;
  const int gi = LaneId() >> lg;
  if (lg >= 3) {
    const int sparse_max = lg >= 4 ? 1 << (lg - 4) : 1;
    const int wid = threadIdx.x / kWave;
    for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < fsize; tile += int64_t(gridDim.x) * kBS) {
      if (threadIdx.x == 0) s_ndense = 0;
      __syncthreads();
      const int64_t i = tile + threadIdx.x;
      if (i < fsize) {
        const int32_t u = F[i];
        const uint64_t pend =
            __hip_atomic_load(reinterpret_cast<unsigned long long*>(&chg[u]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__popcll(pend) <= sparse_max)
          LaneExpand(c, u, off, chunks, S, lg);
        else
          s_dense[atomicAdd(&s_ndense, 1)] = u;
      }
      __syncthreads();
      const int nd = s_ndense;
      for (int j = wid * groups + gi; j < nd; j += (kBS / kWave) * groups) GroupExpand(c, g, s_dense[j], off, chunks, S, lg);
      __syncthreads();
    }
  } else {
    const int64_t wave = (blockIdx.x * int64_t(kBS) + threadIdx.x) / kWave;
    const int64_t waves = int64_t(gridDim.x) * (kBS / kWave);
    for (int64_t base = wave * groups; base < fsize; base += waves * groups) {
      const int64_t i = base + gi;
      if (i >= fsize) continue;  // group-uniform
      GroupExpand(c, g, F[i], off, chunks, S, lg);
    }
  }
  RelaxFlush(c, s_red, s_min);
}

// The hubs' chunks: a group per chunk, each root of the chunk's mask with
// the vertex's current distance (it may have dropped since; a lower
// distance is a true path length and its own pending bit is set).
__global__ __launch_bounds__(kBS) void SbRelaxChunks(const int2* __restrict__ wdst, int32_t* __restrict__ dist,
                                                     uint64_t* __restrict__ chg, int32_t* __restrict__ stamp,
                                                     int32_t* __restrict__ freshstamp, int32_t* __restrict__ qa,
                                                     int32_t* __restrict__ qb, int32_t* __restrict__ fresh,
                                                     const SbChunk* __restrict__ chunks, SbState* S, int lg) {
  if (S->phase != kRelax) return;
  const int64_t nchunks = S->nchunks;
  if (nchunks == 0) return;
  __shared__ LdsQueue<kQueueCap> lnear, lfresh;
  __shared__ unsigned long long s_red[kBS / kWave];
  __shared__ int32_t s_min[kBS / kWave];
  LdsQueueInit(lnear);
  LdsQueueInit(lfresh);
  int32_t* near = S->parity ? qa : qb;
  RelaxCtx c{wdst, dist, chg, stamp, freshstamp, near, fresh, S, &lnear, &lfresh, S->hi, S->near_tag,
             S->fresh_tag, lg, 0, 0, 0, kDistInf};
  const Group g = MakeGroup(lg);
  const int groups = kWave >> lg;
  const int64_t wave = (blockIdx.x * int64_t(kBS) + threadIdx.x) / kWave;
  const int64_t waves = int64_t(gridDim.x) * (kBS / kWave);
  const int gi = LaneId() >> lg;
  for (int64_t base = wave * groups; base < nchunks; base += waves * groups) {
    const int64_t i = base + gi;
    if (i >= nchunks) continue;
    const SbChunk ch = chunks[i];
    const bool active = (ch.mask >> g.r) & 1;
    const int32_t du = active ? LoadC(&dist[(int64_t(ch.u) << lg) + g.r]) : kDistInf;
    RelaxRange(c, g, ch.u, ch.beg, ch.end, active, du);
  }
  RelaxFlush(c, s_red, s_min);
}

// After a relax launch: fold the counters; go on with near, or, when it is
// empty, pick the next bucket from the smallest pending distance (done
// when there is none).  One thread.
__global__ void SbStepRelax(SbState* S) {
  if (threadIdx.x != 0 || S->phase != kRelax) return;
  for (int i = 0; i < kStripes; i++) {
    S->relaxed += static_cast<int64_t>(S->sums[i][0]);
    S->improved += static_cast<int64_t>(S->sums[i][1]);
    S->scanned += static_cast<int64_t>(S->sums[i][2]);
    S->sums[i][0] = S->sums[i][1] = S->sums[i][2] = 0;
  }
  S->relax_launches++;
  S->frontier_in += S->fsize;
  S->nchunks = 0;
  if (S->relax_launches > S->max_launches) {
    S->phase = kFailed;
    return;
  }
  if (S->near_tail > 0) {
    S->parity ^= 1;
    S->fsize = S->near_tail;
    S->near_tail = 0;
    S->near_tag = ++S->tag;
    return;
  }
  S->fsize = 0;
  int32_t m = kDistInf;
  for (int i = 0; i < kStripes; i++) {
    m = S->fresh_min[i] < m ? S->fresh_min[i] : m;
    m = S->pile_min[i] < m ? S->pile_min[i] : m;
  }
  if (m >= kDistInf) {
    S->phase = kDone;
    return;
  }
  S->lo = (static_cast<int64_t>(m) / S->delta) * S->delta;
  S->hi = S->lo + S->delta;
  S->f_tag = ++S->tag;
  S->pile_tag = ++S->tag;
  S->phase = kSplit;
}

// Splits pile p[pile_par] and the fresh pile: entries with a pending
// distance in [lo, hi) join the frontier, those with one at or beyond hi
// the kept pile (each vertex once per split); entries with nothing pending
// are dropped.  A group per entry.
__global__ __launch_bounds__(kBS) void SbSplit(const int32_t* __restrict__ dist, const uint64_t* __restrict__ chg,
                                               int32_t* __restrict__ stamp, int32_t* __restrict__ pilestamp,
                                               int32_t* __restrict__ qa, int32_t* __restrict__ qb,
                                               int32_t* __restrict__ pa, int32_t* __restrict__ pb,
                                               const int32_t* __restrict__ fresh, SbState* S, int lg) {
  if (S->phase != kSplit) return;
  __shared__ LdsQueue<kQueueCap> lf, lp;
  __shared__ int32_t s_min[kBS / kWave];
  LdsQueueInit(lf);
  LdsQueueInit(lp);
  const int32_t* pile = S->pile_par ? pb : pa;
  int32_t* kept = S->pile_par ? pa : pb;
  int32_t* F = S->parity ? qb : qa;
  const int64_t np = S->pile_size, total = np + S->fresh_tail;
  const int64_t hi = S->hi;
  const int32_t f_tag = S->f_tag, pile_tag = S->pile_tag;
  const Group g = MakeGroup(lg);
  const int groups = kWave >> lg;
  const int64_t wave = (blockIdx.x * int64_t(kBS) + threadIdx.x) / kWave;
  const int64_t waves = int64_t(gridDim.x) * (kBS / kWave);
  const int gi = LaneId() >> lg;
  int32_t kmin = kDistInf;
  for (int64_t base = wave * groups; base < total; base += waves * groups) {
    const int64_t i = base + gi;
    bool to_f = false, to_p = false;
    int32_t v = 0;
    if (i < total) {
      v = i < np ? pile[i] : fresh[i - np];
      const uint64_t m = chg[v];
      const bool mine = (m >> g.r) & 1;
      const int32_t d = mine ? dist[(int64_t(v) << lg) + g.r] : kDistInf;
      const bool in_b = mine && d < hi;
      const bool beyond = mine && d >= hi;
      if (beyond && d < kmin) kmin = d;
      const uint64_t fb = GroupBallot(g, in_b), pbits = GroupBallot(g, beyond);
      if (LaneId() == g.leader) {
        if (fb) to_f = atomicExch(&stamp[v], f_tag) != f_tag;
        if (pbits) to_p = atomicExch(&pilestamp[v], pile_tag) != pile_tag;
      }
    }
    LdsQueuePush(lf, to_f, v, F, &S->f_tail);
    LdsQueuePush(lp, to_p, v, kept, &S->pile2_tail);
  }
  LdsQueueFlush(lf, F, &S->f_tail);
  LdsQueueFlush(lp, kept, &S->pile2_tail);
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(kmin, o);
    kmin = t < kmin ? t : kmin;
  }
  if (LaneId() == 0) s_min[threadIdx.x / kWave] = kmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBS / kWave; w++) kmin = s_min[w] < kmin ? s_min[w] : kmin;
    if (kmin < kDistInf) atomicMin(&S->pile_min[blockIdx.x & (kStripes - 1)], kmin);
  }
}

// Clears the minima a split recomputes.  Runs before the split, in the
// split phase only (the pile minimum of the last split stays valid for
// SbStepRelax until then).
__global__ void SbPreSplit(SbState* S) {
  if (S->phase != kSplit) return;
  const int t = threadIdx.x;
  if (t < kStripes) S->pile_min[t] = kDistInf;
}

// After a split: the frontier and the kept pile become current, the fresh
// pile starts over.  An empty frontier (the pile minimum it started from
// was stale) picks the next bucket again at once from the kept pile.
__global__ void SbStepSplit(SbState* S) {
  if (threadIdx.x != 0 || S->phase != kSplit) return;
  S->split_in += S->pile_size + S->fresh_tail;
  S->fsize = S->f_tail;
  S->f_tail = 0;
  S->pile_size = S->pile2_tail;
  S->pile2_tail = 0;
  S->pile_par ^= 1;
  S->fresh_tail = 0;
  S->fresh_tag = ++S->tag;
  for (int i = 0; i < kStripes; i++) S->fresh_min[i] = kDistInf;
  S->buckets++;
  S->near_tag = ++S->tag;
  if (S->fsize > 0) {
    S->phase = kRelax;
    return;
  }
  int32_t m = kDistInf;
  for (int i = 0; i < kStripes; i++) m = S->pile_min[i] < m ? S->pile_min[i] : m;
  if (m >= kDistInf) {
    S->phase = kDone;
    return;
  }
  S->lo = (static_cast<int64_t>(m) / S->delta) * S->delta;
  S->hi = S->lo + S->delta;
  S->f_tag = ++S->tag;
  S->pile_tag = ++S->tag;
}

// dist rows (stride G) -> root-major int32 out[r * n + v] for original ids
// (map: device id of original v, or null).
__global__ __launch_bounds__(kBS) void SbDistOut(const int32_t* __restrict__ dist, int64_t n, int k, int lg,
                                                 const int32_t* __restrict__ map, int32_t* __restrict__ out) {
  const int64_t total = n * k;
  for (int64_t x = blockIdx.x * int64_t(kBS) + threadIdx.x; x < total; x += int64_t(gridDim.x) * kBS) {
    const int64_t r = x / n, v = x % n;
    const int64_t dv = map ? map[v] : v;
    out[x] = dist[(dv << lg) + r];
  }
}

// Reached (root, vertex) pairs and their summed out-degrees.
__global__ __launch_bounds__(kBS) void SbReach(const int32_t* __restrict__ dist, const int64_t* __restrict__ off,
                                               int64_t n, int k, int lg, unsigned long long* out) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long cnt = 0, deg = 0;
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS) {
    int c = 0;
    for (int r = 0; r < k; r++) c += dist[(v << lg) + r] < kDistInf;
    cnt += c;
    deg += static_cast<unsigned long long>(c) * static_cast<unsigned long long>(off[v + 1] - off[v]);
  }
  cnt = BlockSum<kBS>(cnt, s_red);
  deg = BlockSum<kBS>(deg, s_red);
  if (threadIdx.x == 0) {
    if (cnt) atomicAdd(&out[0], cnt);
    if (deg) atomicAdd(&out[1], deg);
  }
}

__global__ void SbZero2(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

struct SbBuffers {
  int32_t* dist;
  uint64_t* chg;
  int32_t *stamp, *freshstamp, *pilestamp;
  int32_t *qa, *qb, *fresh, *pa, *pb;
  SbChunk* chunks;
  SbState* S;
  int lg;
};

// Profile records of one round's kernels (-1 when not profiling).
struct SbLaunched {
  int relax, chunks, split;
};

// Enqueues one round.
SbLaunched SbRound(gg_graph* g, const SbBuffers& b) {
  const int2* wdst = reinterpret_cast<const int2*>(g->out_dst);
  const dim3 blk(kBS), grid(g->cu_count * kBlocksPerCu);
  SbLaunched r;
  r.relax = Run(g, "sssp_batch_relax", 0, SbRelax, grid, blk, static_cast<const int64_t*>(g->out_off), wdst, b.dist, b.chg,
      b.stamp, b.freshstamp, b.qa, b.qb, b.fresh, b.chunks, b.S, b.lg);
  r.chunks = Run(g, "sssp_batch_relax_chunks", 0, SbRelaxChunks, grid, blk, wdst, b.dist, b.chg, b.stamp, b.freshstamp, b.qa,
      b.qb, b.fresh, static_cast<const SbChunk*>(b.chunks), b.S, b.lg);
  Run(g, "sssp_batch_step", 0, SbStepRelax, dim3(1), dim3(64), b.S);
  Run(g, "sssp_batch_step", 0, SbPreSplit, dim3(1), dim3(64), b.S);
  r.split = Run(g, "sssp_batch_split", 0, SbSplit, grid, blk, static_cast<const int32_t*>(b.dist),
      static_cast<const uint64_t*>(b.chg), b.stamp, b.pilestamp, b.qa, b.qb, b.pa, b.pb,
      static_cast<const int32_t*>(b.fresh), b.S, b.lg);
  Run(g, "sssp_batch_step", 0, SbStepSplit, dim3(1), dim3(64), b.S);
  return r;
}

constexpr int kRoundKernels = 6;

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_sssp_batch(gg_graph* g, const int32_t* roots, int k, int32_t delta, int32_t* dist_out,
                             gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(g->weighted, GG_ENOTSUP, "SSSP needs a weighted graph upload");
    GG_REQUIRE(roots != nullptr && k >= 1 && k <= 64, GG_EINVAL, "1 to 64 roots");
    GG_REQUIRE(delta > 0, GG_EINVAL, "delta must be positive");
    for (int i = 0; i < k; i++) GG_REQUIRE(roots[i] >= 0 && roots[i] < g->n, GG_EINVAL, "root out of range");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    hipStream_t st = g->stream;
    int lg = 0;
    while ((1 << lg) < k) lg++;
    SbBuffers b;
    b.lg = lg;
    b.dist = g->Scratch<int32_t>("sb.dist", static_cast<size_t>(n) << lg);
    b.chg = g->Scratch<uint64_t>("sb.chg", n);
    b.stamp = g->Scratch<int32_t>("sb.stamp", n);
    b.freshstamp = g->Scratch<int32_t>("sb.freshstamp", n);
    b.pilestamp = g->Scratch<int32_t>("sb.pilestamp", n);
    // Each list holds a vertex at most once per launch (frontier, near),
    // per split (pile) or between splits (fresh): n entries each.
    b.qa = g->Scratch<int32_t>("sb.qa", n + 64);
    b.qb = g->Scratch<int32_t>("sb.qb", n + 64);
    b.fresh = g->Scratch<int32_t>("sb.fresh", n + 64);
    b.pa = g->Scratch<int32_t>("sb.pa", n + 64);
    b.pb = g->Scratch<int32_t>("sb.pb", n + 64);
    // A vertex yields ceil(deg / kChunkE) <= 2 deg / kSerialMax chunks, once per launch.
    b.chunks = g->Scratch<SbChunk>("sb.chunks", 2 * g->out_entries / kSerialMax + 64);
    b.S = g->Scratch<SbState>("sb.state", 1);
    int32_t* droots = g->Scratch<int32_t>("sb.roots", 64);
    GG_HIP(hipMemcpyAsync(droots, roots, sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    DeviceVertices(g, droots, k);

    const dim3 blk(kBS);
    const double N = static_cast<double>(n);
    int64_t launches = 0;
    BeginTiming(g);
    Run(g, "sssp_batch_init", (4.0 * (1 << lg) + 20) * N, SbInitRows,
        dim3(GridFor(n << lg, kBS, g->cu_count * 32)), blk, b.dist, b.chg, b.stamp, b.freshstamp, b.pilestamp, n,
        lg);
    Run(g, "sssp_batch_init", 0, SbStateInit, dim3(1), dim3(64), b.S, static_cast<int64_t>(delta),
        static_cast<int64_t>(64 * n + 100000));
    Run(g, "sssp_batch_init", 0, SbRoots, dim3(1), dim3(64), static_cast<const int32_t*>(droots), k, lg, b.dist,
        b.chg, b.stamp, b.qa, b.S);
    launches += 3;
    SbState hs;
    auto read_state = [&] {
      GG_HIP(hipMemcpyAsync(&hs, b.S, sizeof(hs), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
    };
    // Bound on replays: every round runs a relax launch or a split.
    const int64_t max_rounds = 2 * (64 * n + 100000) + 16;
    int64_t rounds = 0;
    if (!g->profile) {
      const std::vector<uintptr_t> key = {
          reinterpret_cast<uintptr_t>(b.dist),  reinterpret_cast<uintptr_t>(b.chg),
          reinterpret_cast<uintptr_t>(b.stamp), reinterpret_cast<uintptr_t>(b.freshstamp),
          reinterpret_cast<uintptr_t>(b.pilestamp), reinterpret_cast<uintptr_t>(b.qa),
          reinterpret_cast<uintptr_t>(b.qb),    reinterpret_cast<uintptr_t>(b.fresh),
          reinterpret_cast<uintptr_t>(b.pa),    reinterpret_cast<uintptr_t>(b.pb),
          reinterpret_cast<uintptr_t>(b.chunks), reinterpret_cast<uintptr_t>(b.S),
          reinterpret_cast<uintptr_t>(g->out_off), reinterpret_cast<uintptr_t>(g->out_dst),
          static_cast<uintptr_t>(lg)};
      if (!g->sb_exec || g->sb_exec_key != key) {
        if (g->sb_exec) {
          (void)hipGraphExecDestroy(g->sb_exec);
          g->sb_exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        GG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          for (int i = 0; i < kRoundsPerReplay; i++) SbRound(g, b);
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
        g->sb_exec = exec;
        g->sb_exec_key = key;
      }
      for (;;) {
        GG_HIP(hipGraphLaunch(g->sb_exec, st));
        launches += kRoundKernels * kRoundsPerReplay;
        rounds += kRoundsPerReplay;
        read_state();
        if (hs.phase == kDone || hs.phase == kFailed) break;
        GG_REQUIRE(rounds < max_rounds, GG_EHIP, "batched SSSP did not converge");
      }
    } else {
      // Per-kernel profiling: one round at a time, each launch credited with
      // its algorithmic bytes once the state says what it did.  Relax: per
      // frontier vertex its offsets and pending mask (24 B), per edge a group
      // scanned the {v, w} pair (8 B), per (edge, root) relaxation the
      // neighbour's distance (4 B), per improvement the distance and the
      // pending mask written (12 B).  Split: per pile entry the id, the mask
      // and one distance (16 B).
      SbState prev;
      std::memset(&prev, 0, sizeof(prev));
      for (;;) {
        const SbLaunched r = SbRound(g, b);
        launches += kRoundKernels;
        rounds++;
        read_state();
        CreditBytes(g, r.relax,
                    24.0 * (hs.frontier_in - prev.frontier_in) + 8.0 * (hs.scanned - prev.scanned) +
                        4.0 * (hs.relaxed - prev.relaxed) + 12.0 * (hs.improved - prev.improved));
        CreditBytes(g, r.split, 16.0 * (hs.split_in - prev.split_in));
        prev = hs;
        if (hs.phase == kDone || hs.phase == kFailed) break;
        GG_REQUIRE(rounds < max_rounds, GG_EHIP, "batched SSSP did not converge");
      }
    }
    GG_REQUIRE(hs.phase == kDone, GG_EHIP, "batched SSSP did not converge");
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (dist_out) {
      int32_t* out = g->Scratch<int32_t>("sb.out", static_cast<size_t>(n) * k);
      SbDistOut<<<dim3(GridFor(n * k, kBS, g->cu_count * 16)), blk, 0, st>>>(
          b.dist, n, k, lg, static_cast<const int32_t*>(g->new_id), out);
      GG_HIP(hipGetLastError());
      GG_HIP(hipMemcpyAsync(dist_out, out, sizeof(int32_t) * n * k, hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
    }
    if (stats) {
      unsigned long long* acc = g->Scratch<unsigned long long>("sb.reach", 2);
      SbZero2<<<1, 64, 0, st>>>(acc);
      SbReach<<<dim3(GridFor(n, kBS, g->cu_count * 16)), blk, 0, st>>>(b.dist, g->out_off, n, k, lg, acc);
      unsigned long long h[2];
      GG_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      stats->total_ms = wt.ms();
      stats->iterations = hs.buckets;
      stats->td_steps = hs.relax_launches;
      stats->reached = static_cast<int64_t>(h[0]);
      stats->reached_edges = static_cast<int64_t>(h[1]);
      stats->edges_examined = hs.relaxed;
      stats->vertices_examined = hs.scanned;
      stats->launches = launches;
    }
  });
}
